"""The event timers of the drop-in surface (eventTimer.h:16-32,
eventTimer.cu:6-55) through the C ABI: initEventTimer .. setEventTimer
around a GPU solve on stream 0, their accumulation and unit conversions,
and printEventTimer's output format."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


class EventTimer(C.Structure):
    _fields_ = [("time", C.c_float), ("start", C.c_void_p), ("stop", C.c_void_p)]


def _libc():
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    return libc


@pytest.mark.parametrize("prec", ["s", "d"])
def test_event_timers_around_a_solve(prec, ops4, cuda, tmp_path):
    import torch
    import bbfmmongpu_amd as B
    lib = B.load(prec)
    lib.initEventTimer.argtypes = [C.POINTER(EventTimer)]
    lib.startEventTimer.argtypes = [C.POINTER(EventTimer)]
    lib.stopEventTimer.argtypes = [C.POINTER(EventTimer)]
    lib.finalizeEventTimer.argtypes = [C.POINTER(EventTimer)]
    lib.printEventTimer.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(EventTimer)]
    for f in (lib.initEventTimer, lib.startEventTimer, lib.stopEventTimer, lib.finalizeEventTimer,
              lib.printEventTimer):
        f.restype = None
    lib.readEventTimer.argtypes = [C.POINTER(EventTimer)]
    lib.readEventTimer.restype = C.c_double
    lib.getEventTimer.argtypes = [C.POINTER(EventTimer)]
    lib.getEventTimer.restype = C.c_double
    lib.setEventTimer.argtypes = [C.POINTER(EventTimer), C.c_double]
    lib.setEventTimer.restype = None

    U, K, _ = ops4
    pos, q = O.uniform_inputs(65536, seed=0)
    plan = B.Plan(pos, 4, U, K, prec=prec, stream=0).set(scheme=B.SCHEME_IJ)  # stream 0, as the timers
    dt = torch.float32 if prec == "s" else torch.float64
    qd = torch.from_numpy(q).to(dt).to(cuda)
    phi = plan.solve(qd)
    torch.cuda.synchronize()

    t = EventTimer()
    t.time = 123.0  # initEventTimer resets the accumulator
    lib.initEventTimer(C.byref(t))
    assert lib.bbfmm_amd_last_error() == 0
    assert t.time == 0.0 and t.start and t.stop
    lib.startEventTimer(C.byref(t))
    plan.solve(qd, phi)
    lib.stopEventTimer(C.byref(t))  # synchronises on the stop event
    first = t.time
    assert first > 0.0
    assert lib.readEventTimer(C.byref(t)) == pytest.approx(first, rel=0, abs=0)
    assert lib.getEventTimer(C.byref(t)) == float(np.float32(first) / np.float32(1000))
    # a second interval accumulates
    lib.startEventTimer(C.byref(t))
    plan.solve(qd, phi)
    plan.solve(qd, phi)
    lib.stopEventTimer(C.byref(t))
    assert t.time > first
    assert lib.readEventTimer(C.byref(t)) == float(np.float32(t.time))
    # setEventTimer stores milliseconds as float
    lib.setEventTimer(C.byref(t), 2.5e3 + 1e-4)
    assert t.time == np.float32(2.5e3 + 1e-4)
    assert lib.getEventTimer(C.byref(t)) == float(np.float32(2.5e3 + 1e-4) / np.float32(1000))
    # printEventTimer: "# printEventTimer: <name> = %14.7e" in seconds
    libc = _libc()
    path = str(tmp_path / "timer.txt").encode()
    fp = libc.fopen(path, b"w")
    assert fp
    lib.printEventTimer(fp, b"kernel", C.byref(t))
    libc.fclose(fp)
    text = open(path, "rb").read().decode()
    want = "# printEventTimer: kernel = %14.7e\n" % float(np.float32(t.time) / np.float32(1000))
    assert text == want, (text, want)
    lib.finalizeEventTimer(C.byref(t))
    plan.close()
