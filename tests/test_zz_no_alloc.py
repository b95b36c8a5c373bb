"""md_batch_error_analysis_noalloc, the entry point that allocates nothing,
against md_batch_error_analysis through the C-ABI: two calls in a row, the
second on a smaller replica of the same batch.  The scratch it uses is a
static device array shared by every call on the GPU."""
import numpy as np
import pytest

from md_brushes_amd import MD_FP32, Batch, initialize_position, initialize_velocity, load_library
from md_brushes_amd.abi import _dp
from md_brushes_amd.configs import replica

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def test_noalloc_error_analysis_matches_allocating():
    print("[no-alloc error analysis] two replicas, large then small", flush=True)
    lib = load_library()
    b = Batch(MD_FP32)
    slots = []
    for side, nb, arch, val, ez, seed in ((6, 60, 1, 3, 0.5, 41), (2, 30, 3, 1, 0.2, 42)):
        p = replica(side, nb, arch, val, ez, seed, MD_FP32)
        slots.append((b.add(p, initialize_position(p, seed), initialize_velocity(p, seed + 1)), p.NN))
    b.run(2)
    for slot, nn in slots:
        ref, fref = b.error_analysis(slot)
        out = np.zeros(8)
        f = np.zeros(3 * nn)
        rc = lib.md_batch_error_analysis_noalloc(b.h, slot, _dp(out), _dp(f))
        assert rc == 0, rc
        got = f.reshape(3, nn).T
        print(f"  slot {slot} NN {nn}: rmsf {out[0]:.6e} vs {ref['rmsf']:.6e}, force rel {_rel(got, fref):.1e}")
        assert [int(x) for x in out[5:8]] == [ref["K1"], ref["K2"], ref["K3"]]
        assert out[3] == ref["alpha"] and out[4] == ref["rcc"]
        assert abs(out[1] - ref["f_r"]) <= 1e-12 * abs(ref["f_r"])
        assert abs(out[2] - ref["f_k"]) <= 1e-12 * abs(ref["f_k"])
        assert _rel(got, fref) < 1e-12
        assert abs(out[0] - ref["rmsf"]) < 1e-9 * ref["rmsf"]
    b.close()
