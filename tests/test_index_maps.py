"""Host checks of the lane maps the specialised real-transform kernels use
to find the mirrored half-spectrum bin N - k in another lane's registers
(rocket-fft_amd/csrc/fast.hip: c2r_load_shuffled, r2c_store_shuffled),
for every power-of-two length and radix schedule of that file."""
import pytest

# length -> radix schedule of the specialised kernels (fast.hip go_op)
SCHEDULES = {32: (16, 2), 64: (16, 4), 128: (16, 8), 256: (16, 16), 512: (16, 16, 2),
             1024: (16, 16, 4)}


@pytest.mark.parametrize("n", sorted(SCHEDULES))
def test_c2r_stage0_mirror_map(n):
    """Stage 0: thread j holds inputs p = j + r TPL (R = first radix,
    TPL = n / R); bin n - p is slot R-1-r of lane TPL - j, or for j = 0 slot
    R - r of the lane itself (r >= 1) and bin n itself (r = 0)."""
    R = SCHEDULES[n][0]
    tpl = n // R
    if tpl < 2 or tpl > 64:
        pytest.skip("not a shuffled shape")
    for j in range(tpl):
        for r in range(R):
            p = j + r * tpl
            if j >= 1:
                lane, slot = tpl - j, R - 1 - r
                assert 1 <= lane < tpl
                assert lane + slot * tpl == n - p
            elif r >= 1:
                assert (R - r) * tpl == n - p


@pytest.mark.parametrize("n", sorted(SCHEDULES))
def test_r2c_last_stage_mirror_map(n):
    """Last stage: thread i holds bins k = i + q TPL + r NS (NS = n / R, Q =
    NS / TPL); bin n - k is slot (Q-1-q, R-1-r) of lane TPL - i, or for i = 0
    slot (Q-q, R-1-r) (q >= 1) / (0, R-r) (q = 0, r >= 1) of the lane itself.
    Every bin 0..n-1 is held exactly once."""
    sched = SCHEDULES[n]
    R = sched[-1]
    tpl = n // 16
    ns = n // R
    if tpl < 2 or tpl > 64 or ns % tpl:
        pytest.skip("not a register-store shape")
    Q = ns // tpl
    held = []
    for i in range(tpl):
        for q in range(Q):
            for r in range(R):
                k = i + q * tpl + r * ns
                held.append(k)
                if i >= 1:
                    kp = (tpl - i) + (Q - 1 - q) * tpl + (R - 1 - r) * ns
                elif q >= 1:
                    kp = (Q - q) * tpl + (R - 1 - r) * ns
                elif r >= 1:
                    kp = (R - r) * ns
                else:
                    continue  # bins 0 and n come from z_0 alone
                assert kp == n - k
    assert sorted(held) == list(range(n))
