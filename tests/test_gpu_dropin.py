"""The C++ drop-in headers end to end on the GPU: a program written against
umappp's public surface (initialize with a data matrix and the default
spectral initialisation, then Status::run for all epochs, with a copy of the
Status taken half-way and run to the end beside the original), compiled with
g++ against libumappp_amd, gives a finite embedding whose 15-NN preservation
matches the Python path's on the same input; the copy's embedding equals the
original's bit for bit."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from _data import mixture
from _quality import knn_preservation

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]

PROGRAM = r'''
#include <cstdio>
#include <vector>
#include "umappp/umappp.hpp"

int main(int argc, char** argv) {
    const int n = 3000, dim = 20;
    std::vector<float> data(static_cast<size_t>(n) * dim);
    FILE* f = std::fopen(argv[1], "rb");
    if (!f || std::fread(data.data(), sizeof(float), data.size(), f) != data.size()) return 2;
    std::fclose(f);
    std::vector<float> emb(static_cast<size_t>(n) * 2);
    umappp::Options opt;                       // SPECTRAL initialisation, 500 epochs at this size
    auto status = umappp::initialize(dim, n, data.data(), 2, emb.data(), opt);
    if (status.num_observations() != n || status.num_epochs() != 500) return 3;
    status.run(emb.data(), 250);
    auto copy = status;                        // a deep copy, as the reference's Status copies by value
    std::vector<float> emb2 = emb;
    copy.run(emb2.data());
    if (status.epoch() != 250 || copy.epoch() != 500) return 5;
    status.run(emb.data());
    if (status.epoch() != 500) return 4;
    if (emb != emb2) return 6;                 // the same state gives the same layout, bit for bit
    FILE* o = std::fopen(argv[2], "wb");
    std::fwrite(emb.data(), sizeof(float), emb.size(), o);
    std::fclose(o);
    return 0;
}
'''


def test_cpp_dropin_end_to_end(umx, tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    x, _ = mixture(21, 3000, 20, 8, dtype=np.float32)
    (tmp_path / "data.bin").write_bytes(x.tobytes())
    src = tmp_path / "main.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "main"
    lib_dir = umx.library_path().parent
    subprocess.run(["g++", "-std=c++17", "-O2", f"-I{ROOT / 'include'}", str(src), "-o", str(exe),
                    f"-L{lib_dir}", "-lumappp_amd", f"-Wl,-rpath,{lib_dir}"], check=True)
    r = subprocess.run([str(exe), str(tmp_path / "data.bin"), str(tmp_path / "emb.bin")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    emb = np.frombuffer((tmp_path / "emb.bin").read_bytes(), dtype=np.float32).reshape(3000, 2)
    assert np.all(np.isfinite(emb))
    # the same pipeline through the Python mirror
    idx, dist = umx.find_knn(x, 15)
    e2 = np.zeros((3000, 2), dtype=np.float32)
    st = umx.initialize(umx.NeighborList.from_knn(idx, dist), 2, e2, umx.Options())
    st.run(e2)
    p_cpp = knn_preservation(x, emb, 15, high_idx=idx)
    p_py = knn_preservation(x, e2, 15, high_idx=idx)
    print(f"C++ drop-in preservation {p_cpp:.4f}, Python path {p_py:.4f}")
    assert abs(p_cpp - p_py) <= 0.02
