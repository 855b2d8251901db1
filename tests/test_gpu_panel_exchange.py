"""The QR panels of up to 8 shifts run one shift per XCD and exchange their
column sums through that XCD's L2 when the kernel finds all of a shift's
workgroups there (csrc/qr.hip), else with device-scope stores.  Both forms
sum in the same fixed order, so the polar factor must come out bit for bit
the same with the automatic choice and with device-scope stores forced
(zolo_set_panel_exchange(1), the form the placement check falls back to)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


# n = 700: the factored QR step's 8 stacks of at most 764 rows (256-thread
# panels, slot polling); n = 8320: stacks over 8192 rows as well (512-thread
# panels, flags) and the dense QR of A/alpha over 8192 rows (across XCDs)
@pytest.mark.parametrize("n", [700, 8320])
def test_panel_exchange_modes_agree(torch_gpu, zl, n):
    torch = torch_gpu
    g = torch.Generator(device="cuda")
    g.manual_seed(17 + n)
    dA = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    outs = []
    try:
        for mode in (0, 1):
            zl.set_panel_exchange(mode)
            assert zl.get_panel_exchange() == mode
            dU, _, info = zl.pzolopd_auto_device(dA, "f64", rk=8, polar=False)
            torch.cuda.synchronize()
            outs.append((dU.clone(), info))
            del dU
    finally:
        zl.set_panel_exchange(0)
    (u0, i0), (u1, i1) = outs
    assert i0.schedule == i1.schedule and i0.schedule[0] == "Q", (i0.schedule, i1.schedule)
    assert torch.equal(u0, u1)
    eye = torch.eye(n, dtype=torch.float64, device="cuda")
    assert (torch.linalg.norm(u0.T @ u0 - eye) / np.sqrt(n)).item() < 1e-14


def test_panel_exchange_mode_argument(zl):
    with pytest.raises(Exception):
        zl.set_panel_exchange(2)
    assert zl.get_panel_exchange() == 0
