"""Dense 5-qubit blocks on the matrix cores.  No reference counterpart; the
checks are the dense matrix applied group by group in numpy (exact index
arithmetic), unitarity (norm preserved), and the Kronecker product of five
single-qubit gates against the engine's U3 tile pass (itself pinned to the
oracle)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def haar(rng, d=32):
    """Haar-random unitary: QR of a complex Ginibre matrix with the phases of
    R's diagonal moved into Q (Mezzadri's construction)."""
    z = (rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    ph = np.diag(r) / np.abs(np.diag(r))
    return q * ph[None, :]


def apply_block(psi, n, targets, u):
    """y[t] = sum_t' u[t, t'] x[t'] over the groups of the target bits."""
    tmask = sum(1 << t for t in targets)
    x = np.arange(1 << n, dtype=np.int64)
    base = x[(x & tmask) == 0]
    toff = np.array([sum(((t >> b) & 1) << targets[b] for b in range(5)) for t in range(32)], dtype=np.int64)
    g = base[:, None] | toff[None, :]
    out = psi.astype(np.complex128).copy()
    out[g] = psi.astype(np.complex128)[g] @ u.T
    return out


def rand_state(rng, n, dt):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return (v / np.linalg.norm(v)).astype(dt)


@pytest.mark.parametrize("n,targets", [(10, [0, 1, 2, 3, 4]), (10, [9, 3, 7, 0, 5]), (12, [11, 10, 9, 8, 7]),
                                       (16, [2, 15, 6, 11, 0]), (20, [19, 4, 13, 8, 1]), (9, [8, 1, 4, 2, 6])])
@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_block_matches_dense_groups(gpu_engine, n, targets, prec):
    from zxh_adae_amd import ZXHError
    from zxh_adae_amd.kernels import DeviceState
    rng = np.random.default_rng(n * 31 + targets[0])
    u = haar(rng)
    assert np.max(np.abs(u.conj().T @ u - np.eye(32))) < 1e-12
    dt = np.complex64 if prec == "c64" else np.complex128
    psi = rand_state(rng, n, dt)
    d = DeviceState(n, prec)
    d.upload(psi)
    if prec == "c64" and n < 10:
        with pytest.raises(ZXHError):
            d.block5(targets, u)
        return
    d.block5(targets, u)
    got = d.download().astype(np.complex128)
    want = apply_block(psi, n, targets, u)
    tol = 1e-12 if prec == "c128" else 2e-6
    assert np.max(np.abs(got - want)) < tol
    assert abs(np.linalg.norm(got) - np.linalg.norm(psi.astype(np.complex128))) < (1e-12 if prec == "c128" else 1e-5)


@pytest.mark.parametrize("tiles", [0, 1, 3])
@pytest.mark.parametrize("variant", range(8))
@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_every_kernel_variant_matches_dense_groups(gpu_engine, prec, variant, tiles):
    """Each variant of the block kernel (prefetch, matrix from LDS or
    registers, cached or nontemporal access) and grid (persistent, one or
    three wave tiles per wave), on a state with several wave tiles per wave
    so the prefetch path runs, and low targets."""
    from zxh_adae_amd import autotune
    from zxh_adae_amd.kernels import DeviceState
    knob = f"block.variant.{prec}"
    tknob = f"block.tiles.{prec}"
    old = autotune.tune_get(knob)
    old_tiles = autotune.tune_get(tknob)
    rng = np.random.default_rng(variant + (100 if prec == "c64" else 0))
    n, targets = 23, [0, 1, 7, 15, 22]
    u = haar(rng)
    dt = np.complex64 if prec == "c64" else np.complex128
    psi = rand_state(rng, n, dt)
    d = DeviceState(n, prec)
    d.upload(psi)
    try:
        autotune.tune_set(knob, variant)
        autotune.tune_set(tknob, tiles)
        d.block5(targets, u)
        d.block5(targets[::-1], u.conj().T)
    finally:
        autotune.tune_set(knob, old)
        autotune.tune_set(tknob, old_tiles)
    got = d.download().astype(np.complex128)
    want = apply_block(apply_block(psi, n, targets, u), n, targets[::-1], u.conj().T)
    tol = 1e-12 if prec == "c128" else 4e-6
    assert np.max(np.abs(got - want)) < tol


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_kronecker_block_equals_u3_pass(gpu_engine, prec):
    """u = U4 x U3 x U2 x U1 x U0 (U_b on target b) as one block equals the
    five single-qubit gates applied by the tile pass."""
    from zxh_adae_amd import workloads as W
    from zxh_adae_amd.kernels import DeviceState, pack_sel
    n = 14
    targets = [3, 12, 0, 7, 9]
    rng = np.random.default_rng(4)
    us = [W.u3_matrix(*rng.uniform(0, 2 * np.pi, 3)) for _ in range(5)]
    u = np.array([[1.0 + 0j]])
    for b in range(5):
        u = np.kron(us[b], u)  # bit b of the combination index = target b
    dt = np.complex64 if prec == "c64" else np.complex128
    psi = rand_state(rng, n, dt)
    a = DeviceState(n, prec)
    a.upload(psi)
    a.block5(targets, u)
    b = DeviceState(n, prec)
    b.upload(psi)
    b.u3_batch([(pack_sel(1 << t, False), 1 << t, us[k]) for k, t in enumerate(targets)])
    tol = 1e-12 if prec == "c128" else 3e-6
    assert np.max(np.abs(a.download().astype(np.complex128) - b.download().astype(np.complex128))) < tol


def test_block_rejects_bad_targets(gpu_engine):
    from zxh_adae_amd import ZXHError
    from zxh_adae_amd.kernels import DeviceState
    d = DeviceState(12, "c128")
    u = np.eye(32)
    with pytest.raises(ZXHError):
        d.block5([0, 1, 2, 3, 3], u)
    with pytest.raises(ZXHError):
        d.block5([0, 1, 2, 3, 12], u)


def numpy_layer(psi, n, layer):
    out = psi.astype(np.complex128)
    for targets, u in layer["blocks"]:
        out = apply_block(out, n, targets, u)
    x = np.arange(1 << n, dtype=np.int64)
    th = np.zeros(1 << n)
    for code, cq, q, theta in layer["diag"]:
        if code == 4:
            th += theta * ((x >> q) & 1)
        else:
            th += theta * ((x >> q) & 1) * ((x >> cq) & 1)
    out = out * np.exp(1j * th)
    for mask, t in layer["mcx"]:
        perm = np.where((x & mask) == mask, x ^ (1 << t), x)
        out = out[perm]
    return out


@pytest.mark.parametrize("prec,n", [("c128", 14), ("c64", 15)])
def test_large_random_circuit_layers_and_sampling(gpu_engine, prec, n):
    """Three layers of the block / diagonal / MCX circuit against numpy, then
    4096 shots bit-exact against the oracle's sampler on the GPU's state."""
    import oracle_py as O
    from zxh_adae_amd import workloads as W
    from zxh_adae_amd.kernels import DeviceState, apply_layer, set_basis_zero
    layers = W.c5_layers(n, 3, 55)
    d = DeviceState(n, prec)
    set_basis_zero(d)
    want = np.zeros(1 << n, dtype=np.complex128)
    want[0] = 1.0
    for L in layers:
        apply_layer(d, L)
        want = numpy_layer(want, n, L)
    got = d.download().astype(np.complex128)
    assert np.max(np.abs(got - want)) < (1e-12 if prec == "c128" else 1e-5)
    seed = O.mix_seed(123, 0)
    shots = d.sample(18, seed, 4096)
    ref = O.oracle_sample_state(got, 18, seed, 4096)
    assert np.array_equal(shots, ref)
