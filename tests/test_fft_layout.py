"""CPU test of the LDS layout facts the 16384-point FFT's wave-local third
exchange relies on (sdbe_amd/csrc/sdbe_device.hpp, fft16k_p3w): a host
program built with hipcc against that header checks, with the header's own
lds_re, lds_fold and fft16k_tord,
* fft16k_tord is a permutation of 0..1023;
* in every wave, the words of its pass-3 reads, its pass-3 writes and its
  pass-4 reads are one and the same set of 1024 words, and the 16 waves' sets
  are disjoint, which is what lets the exchange run without a workgroup
  barrier;
* the base-plus-offset word forms written in fft16k_p3w equal
  lds_re(lds_fold(element));
* every one of those instructions reaches 64 distinct banks, and 32 in each
  half-wave.
Runs no GPU code.  A broken fold (bits 10-11 folded in place of 12-13) or the
identity thread order makes the program report failures."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "sdbe_amd", "csrc")

CHECK = r"""
#include <cstdio>
#include <set>
#include <vector>
#include "sdbe_device.hpp"
using namespace sdbe;
static int distinct(const std::vector<int> &w, int lo, int n, int nb) {
    std::set<int> b;
    for (int i = lo; i < lo + n; ++i) b.insert(w[i] % nb);
    return (int)b.size();
}
int main() {
    int bad = 0;
    std::vector<int> seen(1024, 0);
    for (int t = 0; t < 1024; ++t) seen[fft16k_tord(t)]++;
    for (int x : seen) bad += x != 1;
    std::vector<int> owner(2 * NCH, -1);
    for (int w = 0; w < 16; ++w) {
        std::set<int> rd, wr, rd4;
        for (int l = 0; l < 64; ++l) {
            const int t = 64 * w + l, T = fft16k_tord(t), k = T & 255, rt = lds_re(T);
            for (int r = 0; r < 16; ++r) {
                rd.insert(lds_re(T + 1024 * r));
                const int o = 4096 * (T >> 8) + k + 256 * r;
                const int ww = ((rt ^ lds_re(256 * (r & 3))) + 8192 * (T >> 8)) + 2048 * (r >> 2);
                bad += ww != lds_re(lds_fold(o));
                wr.insert(ww);
                const int rr = ((rt ^ lds_re(256 * (r >> 2))) + 8192 * (r >> 2)) + 2048 * (r & 3);
                bad += rr != lds_re(lds_fold(T + 1024 * r));
                rd4.insert(rr);
            }
        }
        bad += !(rd == wr && wr == rd4 && rd.size() == 1024);
        for (int x : rd) {
            bad += owner[x] >= 0;
            owner[x] = w;
        }
        for (int r = 0; r < 16; ++r) {
            std::vector<int> a(64), b(64), c(64);
            for (int l = 0; l < 64; ++l) {
                const int T = fft16k_tord(64 * w + l);
                a[l] = lds_re(T + 1024 * r);
                b[l] = lds_re(lds_fold(4096 * (T >> 8) + (T & 255) + 256 * r));
                c[l] = lds_re(lds_fold(T + 1024 * r));
            }
            for (auto *v : {&a, &b, &c})
                bad += (distinct(*v, 0, 64, 64) != 64) + (distinct(*v, 0, 32, 32) != 32) +
                       (distinct(*v, 32, 32, 32) != 32);
        }
    }
    printf("%s %d\n", bad ? "FAIL" : "OK", bad);
    return bad != 0;
}
"""


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    return None


def test_wave_local_exchange_layout(tmp_path):
    hipcc = _hipcc()
    if hipcc is None:
        pytest.skip("hipcc not found")
    src = tmp_path / "check.cpp"
    src.write_text(CHECK)
    exe = tmp_path / "check"
    subprocess.run([hipcc, "-std=c++17", "--offload-arch=gfx950", "--offload-host-only", "-O1", "-I", CSRC,
                    str(src), "-o", str(exe)], check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr
