"""Simulation output and restart on the GPU path: SimulationManager.simulate(output_directory)
writes HDF5/XDMF files on the reference's save_dt schedule (io_utils/output_writer.py:228-286),
and InitializationManager.initialization(restart_file_path=...) continues from one of them
(material_fields_initializer.py:212-333)."""
import glob
import os

import numpy as np
import pytest

from oracle import jxf_oracle as o

from .test_gpu_parity import _input_manager, say

pytestmark = pytest.mark.gpu


def _sod(end_time, save_dt):
    from jaxfluids_amd import InputManager

    dom, W0 = o.sod_initial(200)
    s = o.Setup(stencil="WENO5-JS", variable="PRIMITIVE", riemann="HLLC", signal_speed="EINFELDT",
                bcs={"east": "ZEROGRADIENT", "west": "ZEROGRADIENT"})
    im = _input_manager(dom, s, end_time=end_time)
    im.case_setup_dict["general"]["save_dt"] = save_dt
    im = InputManager(im.case_setup_dict, im.numerical_setup_dict)
    return dom, W0, im


def test_output_schedule_and_restart(tmp_path):
    from jaxfluids_amd import InitializationManager, SimulationManager
    from jaxfluids_amd import hdf5

    say("[output] Sod 200 cells, save_dt 0.02, t_end 0.1")
    dom, W0, im = _sod(0.1, 0.02)
    init = InitializationManager(im)
    buf = init.initialization(W0)
    sim = SimulationManager(im)
    out = sim.simulate(buf, output_directory=str(tmp_path))
    files = sorted(glob.glob(str(tmp_path / "data_*.h5")), key=lambda p: float(os.path.basename(p)[5:-3]))
    times = [float(os.path.basename(p)[5:-3]) for p in files]
    # t = 0, the first step at or after each multiple of 0.02, and the final time
    assert times[0] == 0.0 and abs(times[-1] - out.time_control_variables.physical_simulation_time) < 1e-9
    for k in range(1, 5):
        assert any(t >= 0.02 * k - 1e-12 and t < 0.02 * k + 0.01 for t in times), (k, times)
    assert len(glob.glob(str(tmp_path / "*.xdmf"))) == len(files) + 1  # + data_time_series.xdmf
    last = hdf5.read(files[-1])
    Wf = out.material_fields.primitives.cpu().numpy()
    assert np.array_equal(last["primitives"]["density"], Wf[0][dom.interior[1:]].T)
    # restart from the middle file and run to the end again: same final state to rounding
    mid = files[len(files) // 2]
    t_mid = float(hdf5.read(mid)["time"])
    say(f"[restart] from t = {t_mid}")
    buf2 = InitializationManager(im).initialization(restart_file_path=mid)
    assert buf2.time_control_variables.physical_simulation_time == t_mid
    out2 = SimulationManager(im).simulate(buf2)
    W2 = out2.material_fields.primitives.cpu().numpy()
    I = dom.interior
    assert np.max(np.abs(W2[I] - Wf[I])) / np.max(np.abs(Wf[I])) < 1e-10
