"""Host test of the workspace growth policy (rocket-fft_amd/csrc/workspace.hpp):
a buffer handed out during graph capture is never freed while the workspace
lives, growth during capture is refused, and buffers never captured are
freed on growth after the stream is synchronised.  The header is compiled
with g++ against a fake allocator that records every call."""
import os
import subprocess
import textwrap

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "rocket-fft_amd", "csrc")

DRIVER = textwrap.dedent(r"""
    #include <cstdio>
    #include <cstdlib>
    #include <set>
    #include <string>
    #include <vector>
    #include "workspace.hpp"

    struct Log {
      std::vector<std::string> ev;
      std::set<void *> live;
      int next = 1;
    };
    static Log g;

    struct FakeMem {
      void *alloc(size_t n) {
        void *p = reinterpret_cast<void *>(size_t(g.next++) << 20);
        g.live.insert(p);
        g.ev.push_back("alloc " + std::to_string(n));
        return p;
      }
      void free(void *p) {
        if (!g.live.erase(p)) { std::printf("FAIL double free\n"); std::exit(1); }
        g.ev.push_back("free");
      }
      void sync() { g.ev.push_back("sync"); }
    };

    #define CHECK(c) do { if (!(c)) { std::printf("FAIL %s line %d\n", #c, __LINE__); std::exit(1); } } while (0)

    int main() {
      using S = rfh::WsSlot<FakeMem>;
      S::Status st;
      {
        S s;
        CHECK(s.get(0, false, &st) == nullptr && st == S::kOk);
        void *a = s.get(100, false, &st);
        CHECK(a && st == S::kOk && s.bytes() == 100);
        // eager growth of a never-captured buffer: sync, then free
        void *b = s.get(120, false, &st);
        CHECK(b && b != a && s.bytes() == 150);  // grows by at least 1.5x
        CHECK(g.live.count(a) == 0 && g.ev.back() == "free" && g.ev[g.ev.size() - 2] == "sync");
        // capture hands out the current buffer and pins it
        void *c = s.get(140, true, &st);
        CHECK(c == b && st == S::kOk && s.pinned());
        // growth during capture is refused, nothing allocated
        const size_t nev = g.ev.size();
        CHECK(s.get(1000, true, &st) == nullptr && st == S::kGrowDuringCapture);
        CHECK(g.ev.size() == nev);
        // eager growth after capture: the pinned buffer is retired, not freed
        void *d = s.get(1000, false, &st);
        CHECK(d && d != b && st == S::kOk && g.live.count(b) == 1 && s.retired() == 1);
        CHECK(!s.pinned());
        // the new buffer is not pinned: its growth frees it again
        void *e = s.get(5000, false, &st);
        CHECK(e && g.live.count(d) == 0 && g.live.count(b) == 1);
        // smaller requests reuse the buffer, captured or not
        CHECK(s.get(10, true, &st) == e && s.get(10, false, &st) == e);
      }
      // the slot's destructor frees the retired and the current buffers
      CHECK(g.live.empty());
      std::printf("OK\n");
      return 0;
    }
""")


def test_workspace_growth_policy(tmp_path):
    src = tmp_path / "ws_policy.cpp"
    exe = tmp_path / "ws_policy"
    src.write_text(DRIVER)
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", f"-I{CSRC}", str(src), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "OK", r.stdout + r.stderr
