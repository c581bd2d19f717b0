"""The ``latec`` command line and the evaluation harness (latec/cli.py,
scripts/eval_tensor.py, scripts/bench_gnn_subset.py): host-side pieces on the
CPU, the round trips on the GPU (their streams against the oracle's)."""

import os
import re

import numpy as np
import pytest

from tests.conftest import has_gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cli_flags_and_defaults():
    from latec_amd.cli import parser

    a = parser().parse_args(["compress", "in.png", "out.msz", "--eb", "2"])
    assert (a.cmd, a.eb, a.rel, a.levels, a.anchor_stride, a.anchor_block, a.radius,
            a.zstd_level, a.eb_ratio, a.tune, a.tune_size_slack, a.predictor) == \
        ("compress", 2.0, False, 4, 16, 1, 1 << 15, 9, None, "fast", 1.05, "interp")
    a = parser().parse_args(["eval", "x.npy", "--eb", "1e-3", "--rel", "--tune", "rd",
                             "--predictor", "interp-linear", "--levels", "3"])
    assert (a.cmd, a.rel, a.tune, a.predictor, a.levels) == ("eval", True, "rd",
                                                             "interp-linear", 3)
    with pytest.raises(SystemExit):
        parser().parse_args(["compress", "a", "b"])  # --eb is required


def test_cap_crop_and_subset_edges():
    from latec_amd.bench_subset import centred_subset
    from latec_amd.eval_tensor import cap_crop

    x = np.arange(40 * 50 * 30, dtype=np.float32).reshape(40, 50, 30)
    assert cap_crop(x, 8, x.size) is x
    c = cap_crop(x, 8, 5000)  # edge int(5000 ** (1/3)) = 17, floored to 16
    assert all(n % 8 == 0 for n in c.shape) and c.size <= 5000
    assert np.array_equal(c, x[(40 - 16) // 2:(40 - 16) // 2 + 16, 17:33, 7:23])
    s = centred_subset(x, 24, 8)
    assert s.shape == (24, 24, 24) and s[0, 0, 0] == x[8, 13, 3]


def test_report_line_format():
    from latec_amd.eval_tensor import report_line

    x = np.linspace(0, 1, 1000, dtype=np.float32).reshape(10, 100)
    r = x + np.float32(1e-4)
    line, ok = report_line("interp", x, r, 500, 1e-3, 0.12, 0.05)
    assert ok
    assert re.fullmatch(r"  \[interp\]        500 B  ratio    8\.00    4\.000 bpv  "
                        r"PSNR  80\.00 dB  maxerr 0\.0001 PASS  0\.1s/0\.1s", line), line
    _, ok = report_line("gnn", x, x + np.float32(2e-3), 500, 1e-3)
    assert not ok


@pytest.mark.gpu
def test_gpu_cli_image_round_trip(tmp_path, capsys):
    if not has_gpu():
        pytest.skip("no GPU")
    from PIL import Image

    from latec_amd.cli import main
    from oracle import oracle as O

    rng = np.random.RandomState(3)
    img = np.cumsum(np.cumsum(rng.randn(96, 80, 3), 0), 1)
    img = ((img - img.min()) / (img.max() - img.min()) * 255).astype(np.uint8)
    src, msz, rec = tmp_path / "in.png", tmp_path / "a.msz", tmp_path / "rec.png"
    Image.fromarray(img).save(src)
    main(["compress", str(src), str(msz), "--eb", "2", "--levels", "3",
          "--anchor-stride", "8"])
    out = capsys.readouterr().out
    assert "bpp" in out and "outliers" in out
    stream = msz.read_bytes()
    assert stream == O.compress(img, 2.0, "cubic", 3, 8, 1, None, 1)[0]
    main(["decompress", str(msz), str(rec)])
    back = np.asarray(Image.open(rec))
    assert back.shape == img.shape and np.abs(back.astype(int) - img.astype(int)).max() <= 2
    main(["eval", str(src), "--eb", "2", "--levels", "3", "--anchor-stride", "8"])
    out = capsys.readouterr().out
    assert "max abs err" in out and "PASS" in out and "zstd-raw" in out


@pytest.mark.gpu
def test_gpu_eval_tensor_and_subset_bench(tmp_path, capsys):
    if not has_gpu():
        pytest.skip("no GPU")
    from latec_amd import bench_subset, eval_tensor

    rng = np.random.RandomState(4)
    x = np.cumsum(np.cumsum(rng.randn(24, 20, 16, 16), 0), 3).astype(np.float64)
    path = tmp_path / "field.npy"
    np.save(path, x)
    assert eval_tensor.main([str(path), "--eb", "1e-3", "--rel", "--levels", "3",
                             "--anchor-stride", "8"])
    out = capsys.readouterr().out
    assert "float64 input cast to float32" in out and "[interp]" in out and "PASS" in out
    assert eval_tensor.main([str(path), "--eb", "1e-3", "--rel", "--predictor", "gnn",
                             "--levels", "3", "--anchor-stride", "8"])
    out = capsys.readouterr().out
    assert "[gnn   ]" in out and "[interp]" in out and out.count("PASS") == 2
    bench_subset.main([str(path), "--eb", "1e-3", "--rel", "--levels", "3",
                       "--subset-edge", "16", "--label", "t"])
    out = capsys.readouterr().out
    for key in ("PSNR", "max error", "bits/value", "Mvox/s", "peak host RAM",
                "peak GPU alloc/resv"):
        assert key in out, key
    assert "PASS" in out
