"""Host test of the per-line slot addressing of line groups
(rocket-fft_amd/csrc/lines.hpp), the code every line kernel uses to find
where the lines of its group start -- including partial groups, the class of
an earlier GPU fault (a Bluestein row pass of 252 lines in groups of 16 read
the slot of a missing line).  Compiled with g++ and checked on many seeded
line sets against an independent multi-index computation:

* every existing line's input/output offsets and aux index are right;
* the slots of missing lines (past the end of the set) hold zeros;
* over all groups every line of the set appears exactly once;
* every offset stays inside the set's span."""
import os
import subprocess
import textwrap

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "rocket-fft_amd", "csrc")

DRIVER = textwrap.dedent(r"""
    #include <cstdio>
    #include <cstdlib>
    #include <random>
    #include <vector>
    #include "lines.hpp"
    using namespace rfh;
    #define CHECK(c) do { if (!(c)) { std::printf("FAIL %s (case %d)\n", #c, cs); return 1; } } while (0)

    int main() {
      std::mt19937_64 rng(7);
      for (int cs = 0; cs < 4000; ++cs) {
        int nb = 1 + int(rng() % 4);
        int64_t bdim[8], bsi[8], bso[8];
        int64_t n = 1;
        for (int d = 0; d < nb; ++d) {
          bdim[d] = 1 + int64_t(rng() % 9);
          n *= bdim[d];
        }
        // strides: row-major over the batch dims times a line pitch, plus
        // random signs (reversed views) on the input side
        int64_t si = 3 + int64_t(rng() % 5), so = 2 + int64_t(rng() % 7);
        for (int d = nb - 1; d >= 0; --d) {
          bsi[d] = (rng() % 4 == 0 ? -1 : 1) * si;
          bso[d] = so;
          si *= bdim[d];
          so *= bdim[d];
        }
        if (cs == 0) {  // the faulting case: 252 lines (7 x 36), groups of 16
          nb = 2; bdim[0] = 7; bdim[1] = 36; n = 252;
          bsi[1] = 2; bso[1] = 2; bsi[0] = 72; bso[0] = 72;
        }
        const int auxdim = int(rng() % nb);
        const int L = 1 << int(rng() % 6);
        std::vector<int> seen(size_t(n), 0);
        int64_t span_i = 0, span_o = 0;
        for (int d = 0; d < nb; ++d) {
          span_i += (bdim[d] - 1) * (bsi[d] < 0 ? -bsi[d] : bsi[d]);
          span_o += (bdim[d] - 1) * bso[d];
        }
        for (int64_t b0 = 0; b0 < n; b0 += L) {
          const int nl = group_lines(b0, L, n);
          CHECK(nl >= 1 && nl <= L && (nl == L || b0 + nl == n));
          for (int l = 0; l < L; ++l) {
            const LineOffsets o = group_slot(b0, l, L, n, nb, bdim, bsi, bso, auxdim);
            if (l >= nl) {
              CHECK(o.oi == 0 && o.oo == 0 && o.aux == 0);
              continue;
            }
            const int64_t b = b0 + l;
            // independent: multi-index by successive division from the front
            int64_t idx[8], rest = b, stride = n;
            for (int d = 0; d < nb; ++d) {
              stride /= bdim[d];
              idx[d] = rest / stride;
              rest %= stride;
            }
            int64_t ei = 0, eo = 0;
            for (int d = 0; d < nb; ++d) {
              ei += idx[d] * bsi[d];
              eo += idx[d] * bso[d];
            }
            CHECK(o.oi == ei && o.oo == eo && o.aux == int(idx[auxdim]));
            CHECK((o.oi < 0 ? -o.oi : o.oi) <= span_i && o.oo >= 0 && o.oo <= span_o);
            ++seen[size_t(b)];
          }
        }
        for (int64_t b = 0; b < n; ++b) CHECK(seen[size_t(b)] == 1);
      }
      std::printf("OK\n");
      return 0;
    }
""")


def test_line_group_slots(tmp_path):
    src = tmp_path / "slots.cpp"
    exe = tmp_path / "slots"
    src.write_text(DRIVER)
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", f"-I{CSRC}", str(src), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "OK", r.stdout + r.stderr
