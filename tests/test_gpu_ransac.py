"""RANSAC hypothesis scoring on the GPU: the scoring kernel against the
oracle's Sampson classification, bit for bit, and the whole model search
(wslam_ransac: hypotheses fitted on the host, scored a round at a time on
the GPU) against the oracle's restatement of RansacCPU::execute."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import ROOT
from tests.ransac_data import normalized, two_view
from uni_slam_amd import orbfe

pytestmark = pytest.mark.gpu
PKG = os.path.join(ROOT, "uni_slam_amd")


def _say(*a):
    print(*a, flush=True)


def test_score_kernel_equals_oracle(torch_gpu):
    torch = torch_gpu
    p, c, inl, F = two_view(11, 5000, 0.3, 0.7)
    rng = np.random.default_rng(12)
    hyps = [F, F * 2.0 ** 30, np.zeros((3, 3)),
            np.array([[0, 0, 0], [0, 0, -1], [0, 1, 0]], np.float64)]
    for _ in range(296):  # fits of random 8-samples: models near and far from the truth
        idx = rng.choice(len(p), 8, replace=False)
        f = O.fit_fundamental(p[idx], c[idx])
        hyps.append(f if f is not None else rng.normal(size=(3, 3)))
    Fs = np.ascontiguousarray(np.stack(hyps).reshape(-1, 9))
    k, n = Fs.shape[0], len(p)
    want_c, want_m = O.sampson_counts(Fs, p, c, 3.84, with_mask=True)
    with orbfe.Context(64, 48) as ctx:
        dF = torch.from_numpy(Fs).cuda()
        dp = torch.from_numpy(np.ascontiguousarray(p)).cuda()
        dc = torch.from_numpy(np.ascontiguousarray(c)).cuda()
        counts = torch.full((k,), 7, dtype=torch.int32, device="cuda")
        mask = torch.full((k * n,), 9, dtype=torch.uint8, device="cuda")
        _say(f"ransac score kernel {k} hypotheses x {n}")
        ctx.ransac_score_async(dF, k, dp, dc, n, 3.84, counts, mask)
        ctx.ransac_score_async(dF, k, dp, dc, n, 3.84, counts)  # counts only
        torch.cuda.synchronize()
        got_c = counts.cpu().numpy().view(np.uint32)
        got_m = mask.cpu().numpy().reshape(k, n)
    assert np.array_equal(got_c, want_c)
    assert np.array_equal(got_m, want_m)
    assert want_c[0] > 0.6 * n and want_c[2] == 0


def _run_driver(tmp_path, sets, rnd):
    path = tmp_path / "sets.bin"
    with open(path, "wb") as f:
        f.write(np.uint32(len(sets)).tobytes())
        for p, c in sets:
            f.write(np.uint32(len(p)).tobytes())
            f.write(np.ascontiguousarray(p, np.float64).tobytes())
            f.write(np.ascontiguousarray(c, np.float64).tobytes())
    _say(f"wslam_ransac {len(sets)} sets, round {rnd}")
    r = subprocess.run([os.path.join(PKG, "wslam_ransac"), str(path), str(rnd)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    out = []
    for i in range(0, len(lines), 3):
        h = lines[i].split()
        st = {"total": int(h[3]), "iterations": int(h[5]), "inliers": int(h[7]),
              "found": h[9] == "1"}
        idx = np.array([int(v) for v in lines[i + 1].split()[1:]], np.int64)
        F = np.array([float(v) for v in lines[i + 2].split()[1:]]).reshape(3, 3)
        out.append((st, idx, F))
    return out


def test_model_search_equals_oracle(tmp_path, torch_gpu):
    """Three consecutive sets with one generator (as frames are): exact
    two-view data with 30% outliers, 0.5-px noise with 40% outliers, and a
    set too small to fit.  Whatever the round size (1: the reference's one
    hypothesis per step; 5 and 64: batched launches with the generator
    rewound where the adaptive bound stops), the iterations run and the
    inlier set equal the oracle's; the model agrees to rounding."""
    data = [two_view(21, 400, 0.3), two_view(22, 800, 0.4, 0.5), two_view(23, 6, 0.0)]
    sets = [(p, c) for p, c, _, _ in data]
    want = O.ransac(sets)
    for rnd in (1, 5, 64):
        got = _run_driver(tmp_path, sets, rnd)
        assert len(got) == 3
        for s, ((st, idx, F), (wst, wF, wm)) in enumerate(zip(got, want)):
            assert st == wst, (rnd, s)
            assert np.array_equal(idx, np.nonzero(wm)[0]), (rnd, s)
            if wst["found"]:
                assert np.abs(normalized(F) - normalized(wF)).max() < 1e-6, (rnd, s)
    assert want[0][0]["found"] and want[1][0]["found"] and not want[2][0]["found"]
