"""sp::MatFile (include/sp/MatFile.hpp), the MAT v5 writer mirroring the reference's
src/MatFile.h:10-168 / src/MatFile.cpp:10-76: files written by a small C++ program are
read back with scipy.io.loadmat, an independent reader of the public format."""
import os
import subprocess

import numpy as np
import pytest

sio = pytest.importorskip("scipy.io")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROG = r"""
#include <vector>
#include "sp/MatFile.hpp"
struct Dom { int i, j, k; int iSize() const { return i; } int jSize() const { return j; } int kSize() const { return k; } };
struct HostField {  // the field interface MatFile::addField uses
    Dom d; std::vector<double> v; std::string n;
    const Dom& calculationDomain() const { return d; }
    double operator()(int i, int j, int k) const { return v[i + d.i * (j + d.j * k)]; }
    void SynchronizeHostStorage() const {}
    const std::string& name() const { return n; }
};
int main(int argc, char** argv) {
    const int ni = 5, nj = 3, nk = 4;
    std::vector<double> a(ni * nj * nk);
    for (int k = 0; k < nk; ++k) for (int j = 0; j < nj; ++j) for (int i = 0; i < ni; ++i)
        a[i + ni * (j + nj * k)] = i + 100. * j + 10000. * k + 0.25;
    HostField f{{ni, nj, nk}, a, "qreference"};
    {
        sp::MatFile m(argv[1]);
        m.addArray("q", a.data(), ni, nj, nk);
        m.addField(f);
        m.addField("step", f, 7);
    }
    {
        sp::MatFile m(argv[2]);
        m.startCell("q", 2);
        m.addField(f, 0);
        m.addArray("q", a.data(), ni, nj, nk);
        m.endCell();
    }
    return 0;
}
"""


@pytest.fixture(scope="module")
def written(tmp_path_factory):
    d = tmp_path_factory.mktemp("mat")
    src = d / "w.cpp"
    src.write_text(PROG)
    exe = d / "w"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    a, b = d / "a.mat", d / "b.mat"
    subprocess.run([str(exe), str(a), str(b)], check=True)
    k, j, i = np.meshgrid(np.arange(4), np.arange(3), np.arange(5), indexing="ij")
    want = (i + 100. * j + 10000. * k + 0.25).transpose(2, 1, 0)  # MATLAB dims [i j k]
    return a, b, want


def test_named_arrays_round_trip(written):
    a, _, want = written
    m = sio.loadmat(str(a))
    for name in ("q", "qreference", "step_7"):
        assert m[name].shape == (5, 3, 4)
        assert m[name].dtype == np.float64
        assert np.array_equal(m[name], want)


def test_header_and_cell(written):
    a, b, want = written
    head = open(a, "rb").read(128)
    assert head.startswith(b"MATLAB 5.0 MAT-file")
    assert head[124:126] == b"\x00\x01" and head[126:128] == b"IM"
    c = sio.loadmat(str(b))["q"]
    assert c.shape == (2, 1)
    assert np.array_equal(c[0, 0], want) and np.array_equal(c[1, 0], want)
