"""rigc_narrow_phase_enqueue (device-resident inputs and output, kernels
only) captured in a HIP graph through torch's graph capture on a side
stream, and replayed on new contents of the same device buffers: the eager
call and every replay return rigc_narrow_phase's earliest TOI for the same
inputs on a second context, bit for bit, with eager torch work and
synchronisations between the replays."""
import numpy as np
import pytest

from gpu_util import context
from test_gpu_broadphase import cluster, moved

pytestmark = pytest.mark.gpu


def test_enqueue_matches_narrow_phase_and_replays_in_a_graph():
    torch = pytest.importorskip("torch")
    sc = cluster(11, n=27)
    p0, p1 = moved(sc, 71, dpos=0.05, drot=0.1)
    ref = context()
    ref.set_scene(sc)
    ctx = context()
    ctx.set_scene(sc)
    ee, fv = ref.broad_phase(p0, p1, 0.0)
    ee = np.ascontiguousarray(np.asarray(ee, np.int32).reshape(-1, 2))
    fv = np.ascontiguousarray(np.asarray(fv, np.int32).reshape(-1, 2))
    print(f"{len(ee)} edge-edge and {len(fv)} face-vertex candidates", flush=True)
    dev = torch.device("cuda:0")
    P0 = torch.tensor(p0, dtype=torch.float64, device=dev).contiguous()
    P1 = torch.tensor(p1, dtype=torch.float64, device=dev).contiguous()
    EE = torch.tensor(ee, device=dev).contiguous()
    FV = torch.tensor(fv, device=dev).contiguous()
    OUT = torch.full((1,), -1.0, dtype=torch.float64, device=dev)
    s = torch.cuda.Stream()

    def enqueue():
        ctx.narrow_phase_enqueue(s.cuda_stream, P0.data_ptr(), P1.data_ptr(), EE.data_ptr(), len(ee), FV.data_ptr(),
                                 len(fv), OUT.data_ptr())

    want = ref.narrow_phase(p0, p1, ee, fv)[0]
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        enqueue()  # also sizes the context's state outside capture
    s.synchronize()
    print("eager", OUT.item(), "want", want, flush=True)
    assert OUT.item() == want
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        enqueue()
    torch.cuda.synchronize()
    finite = 0
    for k, scale in enumerate((1.0, 2.0, 1.5, 1.0)):
        p1k = p0 + scale * (p1 - p0)
        P1.copy_(torch.tensor(p1k, dtype=torch.float64, device=dev))
        OUT.fill_(-1.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        want = ref.narrow_phase(p0, p1k, ee, fv)[0]
        print(f"replay {k} (scale {scale}): {OUT.item()} want {want}", flush=True)
        assert OUT.item() == want
        finite += int(np.isfinite(want))
    assert finite == 4
    del g
    ctx.close()
    ref.close()
