"""Graph replay of the major cycle (chip_vset_cycle_graph) on the GPU.

A W-projection cube's residual refresh is captured once into a HIP graph and
replayed; between replays the test runs its own eager work on the model
buffer (torch copies, which change the model in place so that the baked-in
pointers stay valid) and synchronises.  Every replay is checked against the
eager major cycle of the same model to 1e-5 of the residual peak (the binned
kernel's float atomics add in a different order from run to run, and the
grid correction amplifies that rounding up to ~25x per axis at the image
corners; nothing else differs).  Replays are counted by the library, so the test also shows that the
graph ran rather than the eager path.
"""
import numpy as np
import pytest
import torch

import synth

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def t(x):
    return torch.from_numpy(np.ascontiguousarray(x))


def cube(gpu, n_chan=4, nx=96, planes=8):
    enu = synth.y_array(3, arm_m=2000.0)
    uvw = synth.uvw_tracks(enu, 34.0, 30.0, 0.5, 60.0)
    n_rows = uvw.shape[0]
    freq = np.linspace(1.0e9, 1.6e9, n_chan)
    scale = freq / synth.C
    rng = np.random.default_rng(23)
    vis = (rng.normal(size=(n_rows, n_chan)) + 1j * rng.normal(size=(n_rows, n_chan))).astype(np.complex64)
    wt = rng.uniform(0.5, 2.0, (n_rows, n_chan)).astype(np.float32)
    flag = rng.uniform(size=(n_rows, n_chan)) < 0.2
    maxuv = np.max(np.abs(uvw[:, :2])) * scale[-1]
    cell = 1.0 / (3.0 * maxuv)
    maxw = np.max(np.abs(uvw[:, 2])) * scale[-1]
    g = gpu.StandardGridder(nx, nx, cell, cell)
    g.attach_wproject(maxw, planes=planes)
    batch = gpu.RowChannelBatch(t(uvw), t(scale), t(vis), t(wt), t(flag.astype(np.uint8)))
    vs = g.visibility_set(batch)
    psf = g.new_grids(n_chan)
    vs.grid(gpu.MODE_PSF, psf=psf)
    nsum = np.array([s["normalization_sumwt"] for s in vs.sums])
    _, _, peaks = g.finish_products(psf, None, nsum)
    return g, vs, nsum, peaks, batch


def models(seed, n_chan, nx, n=20):
    rng = np.random.default_rng(seed)
    m = np.zeros((n_chan, nx, nx), np.float32)
    for c in range(n_chan):
        m[c, rng.integers(15, nx - 15, n), rng.integers(15, nx - 15, n)] = rng.uniform(0.1, 1.0, n)
    return m


def test_major_cycle_graph_replays_match_eager(gpu):
    n_chan, nx = 4, 96
    g, vs, nsum, peaks, batch = cube(gpu, n_chan, nx)
    dev = torch.device("cuda")
    sources = [t(models(s, n_chan, nx)).to(dev) for s in (31, 32, 33)]
    # eager results, one per model
    eager = [vs.major_cycle(m, nsum, peaks).cpu().numpy() for m in sources]
    assert vs.cycle_replays() == 0

    vs.set_cycle_graph(True)
    model = torch.empty_like(sources[0])
    out = torch.empty((n_chan, nx, nx), dtype=torch.float32, device=dev)
    for k in range(6):
        j = k % len(sources)
        model.copy_(sources[j])          # eager work between replays
        scratch = torch.randn(1 << 20, device=dev)
        torch.cuda.synchronize()
        vs.major_cycle(model, nsum, peaks, out=out)
        got = out.cpu().numpy()
        print(f"replay {k}: {rel(got, eager[j]):.2e} of peak")
        assert rel(got, eager[j]) < 1e-5
        del scratch
    # the first call captured and launched, every later call replayed
    assert vs.cycle_replays() == 6
    assert vs.stats()["last_kernel_ms"] < 0  # no binned-kernel events inside the graph

    # new pointers capture again; eager mode afterwards still agrees
    other = sources[1].clone()
    res = vs.major_cycle(other, nsum, peaks).cpu().numpy()
    assert rel(res, eager[1]) < 1e-5
    assert vs.cycle_replays() == 7
    vs.set_cycle_graph(False)
    res = vs.major_cycle(sources[2], nsum, peaks).cpu().numpy()
    assert rel(res, eager[2]) < 1e-5
    assert vs.cycle_replays() == 7
    assert vs.stats()["last_kernel_ms"] > 0


def test_graph_needs_fused_transforms_and_resident_tables(gpu):
    n_chan, nx = 2, 96
    g, vs, nsum, peaks, batch = cube(gpu, n_chan, nx)
    vs.set_cycle_strategy(gpu.CYCLE_HIPFFT)
    vs.set_cycle_graph(True)
    m = t(models(5, n_chan, nx))
    with pytest.raises(gpu.InvalidRequest, match="fused"):
        vs.major_cycle(m, nsum, peaks)
    # W tables recomputed on the fly cannot be captured
    assert g.set_wkernel_budget(1) == 1  # below one plane's tables: one plane per chunk
    vs_otf = g.visibility_set(batch)
    with pytest.raises(gpu.InvalidRequest, match="on the fly"):
        vs_otf.set_cycle_graph(True)
