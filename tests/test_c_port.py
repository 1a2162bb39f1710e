"""The C/OpenMP restatement (CPU baseline) agrees with the NumPy oracle."""
import numpy as np
import pytest

from oracle import jxf_oracle as o
from oracle import c_port


@pytest.mark.parametrize("var,stencil,rs,ss,bc", [
    ("PRIMITIVE", "WENO5-JS", "HLLC", "EINFELDT", "ZEROGRADIENT"),
    ("CHAR-PRIMITIVE", "WENO5-Z", "HLLC", "EINFELDT", "PERIODIC"),
    ("CONSERVATIVE", "WENO5-Z", "RUSANOV", "EINFELDT", "SYMMETRY"),
    ("CHAR-PRIMITIVE", "WENO5-JS", "HLLC", "TORO", "PERIODIC"),
])
def test_c_port_matches_oracle(var, stencil, rs, ss, bc):
    dom, W0 = o.tgv_initial(12, nh=4)
    rng = np.random.default_rng(3)
    W0 = W0 * (1 + 1e-3 * rng.standard_normal(W0.shape))
    s = o.Setup(stencil=stencil, variable=var, riemann=rs, signal_speed=ss, bcs={f: bc for f in o.FACE_LOCATIONS})
    W, U = o.create_buffers(W0, dom, s)
    st = c_port.CPUStepper(dom, s)
    dt = o.compute_dt(W, dom, s)
    assert abs(st.dt(W, o.EPS64) - dt) <= 1e-15 * dt
    Wr, Ur = o.rk3_step(W, U, dt, dom, s)
    Wc, Uc = W.copy(), U.copy()
    st.step(Wc, Uc, dt)
    I = dom.interior
    err = np.max(np.abs(Wc[I] - Wr[I])) / np.max(np.abs(Wr[I]))
    assert err < 1e-12, err
