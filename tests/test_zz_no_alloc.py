"""ocfd_chem_adaptive_batch_noalloc, the chemistry batch entry that allocates
nothing (its workspace is a static device array), against
ocfd_chem_adaptive_batch with the same scheduler (1, the per-lane work queue)
through the C-ABI.  Both run the same per-cell arithmetic, so the bar is bit
for bit: partial densities, energies and accepted substeps per cell.  Two
calls in a row, the second on a smaller batch (the workspace keeps the first
call's contents), and the size limit of the workspace."""
import numpy as np
import pytest

from conftest import load_package, gpu_available
import chemcase

pytestmark = pytest.mark.gpu

MAX_CELLS = 1 << 20


@pytest.fixture(scope="module")
def env():
    if not gpu_available():
        pytest.skip("no GPU")
    pkg = load_package()
    from ocfd_amd import chemistry
    b, M, orc = chemcase.jet_batch(1000)
    return pkg, chemistry.ChemBatch(M), b


def batch(b, n, seed):
    """n cells tiled from the 1000-cell jet batch, dt scaled by 10^U[0,2) so
    the substep counts spread over two decades and the cell order matters."""
    rng = np.random.default_rng(seed)
    idx = np.arange(n) % b["E"].shape[0]
    rho = np.ascontiguousarray(b["rho"][:, idx])
    E = b["E"][idx].copy()
    dt = b["dt"][idx] * 10.0 ** rng.uniform(0.0, 2.0, n)
    return rho, E, dt


def run(pkg, chem, rho, E, dt, noalloc):
    import torch
    r = torch.tensor(rho, dtype=torch.float64, device="cuda").contiguous()
    e = torch.tensor(E, dtype=torch.float64, device="cuda").contiguous()
    d = torch.tensor(dt, dtype=torch.float64, device="cuda").contiguous()
    ns = torch.full((E.shape[0],), -1, dtype=torch.int32, device="cuda")
    n = chem._args(r, e, d)  # checks shapes and sets torch's current stream
    ctl = (0.05, 1e-6, 0.01, 3000)
    if noalloc:
        rc = chem.lib.ocfd_chem_adaptive_batch_noalloc(n, r.data_ptr(), e.data_ptr(), d.data_ptr(), *ctl,
                                                       ns.data_ptr())
    else:
        rc = chem.lib.ocfd_chem_adaptive_batch(n, r.data_ptr(), e.data_ptr(), d.data_ptr(), *ctl, ns.data_ptr(), 1)
    chem._check(rc)
    torch.cuda.synchronize()
    return r.cpu().numpy(), e.cpu().numpy(), ns.cpu().numpy()


def test_noalloc_matches_allocating_entry(env):
    pkg, chem, b = env
    for call, (n, seed) in enumerate(((65536, 11), (5000, 12))):
        rho, E, dt = batch(b, n, seed)
        ref = run(pkg, chem, rho, E, dt, noalloc=False)
        got = run(pkg, chem, rho, E, dt, noalloc=True)
        print(f"call {call + 1}: n = {n}, substeps mean {ref[2].mean():.1f} max {ref[2].max()}")
        assert ref[2].min() >= 1
        for name, a, c in zip(("rho", "E", "nsub"), ref, got):
            assert np.array_equal(a, c), f"call {call + 1} (n = {n}): {name} differs"


def test_noalloc_size_limit(env):
    pkg, chem, b = env
    import torch
    r = torch.zeros((9, 1), dtype=torch.float64, device="cuda")
    e = torch.zeros(1, dtype=torch.float64, device="cuda")
    rc = chem.lib.ocfd_chem_adaptive_batch_noalloc(MAX_CELLS + 1, r.data_ptr(), e.data_ptr(), e.data_ptr(),
                                                   0.05, 1e-6, 0.01, 100, None)
    assert rc != 0
    assert "2^20" in chem.lib.ocfd_chem_last_error().decode()
    # an empty batch is accepted and touches nothing
    rc = chem.lib.ocfd_chem_adaptive_batch_noalloc(0, r.data_ptr(), e.data_ptr(), e.data_ptr(),
                                                   0.05, 1e-6, 0.01, 100, None)
    assert rc == 0
