"""Fork contract (host test, no GPU compute).

The reference engine survives fork(): its thread pool is rebuilt in the child
by pthread_atfork handlers (rocket_fft/_pocketfft_hdronly.h:988-1009;
tests/test_multiprocessing.py:15-35).  HIP state does not survive fork(), so
this library records the process of its first call and, in a child forked
after that, fails every entry point cleanly (error code 4, a message naming
the parent) before making any HIP call -- no hang, no crash.  A child forked
before the parent's first call is unaffected (the library initialises HIP
lazily)."""
import os
import sys

import numpy as np
import pytest

import rocket_fft_amd as R


def _child(fn):
    pid = os.fork()
    if pid == 0:  # child: report through the exit status only
        code = 1
        try:
            code = fn()
        except BaseException:
            code = 3
        finally:
            os._exit(code)
    _, status = os.waitpid(pid, 0)
    return os.waitstatus_to_exitcode(status)


@pytest.mark.skipif(sys.platform != "linux", reason="fork")
def test_forked_child_fails_cleanly():
    x = np.ones((2, 8), np.complex128)
    y = np.zeros_like(x)
    try:  # the parent's first call (the GPU may be absent here: any outcome)
        R.c2c(x, y, np.array([1]), True, 1.0, 1)
    except RuntimeError:
        pass

    def in_child():
        try:
            R.c2c(x, y, np.array([1]), True, 1.0, 1)
        except RuntimeError as e:
            code, _ = R.last_error()
            if code == 4 and f"forked from process {os.getppid()}" in str(e):
                return 0
            return 2
        return 4

    assert _child(in_child) == 0
    # the parent is unaffected by the child's attempt
    code, _ = R.last_error()
    assert code != 4
