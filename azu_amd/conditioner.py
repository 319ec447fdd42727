"""Host mirror of the depth half of kfusion::sensor::SignalConditioner
(include/sensor/SignalConditioner.h:29-47, src/sensor/SignalConditioner_omp.cpp:337-716)
over the gfx950 kernels of the C ABI (azu_cond_*).

process() runs the reference's depth stages in its order on raw Kinect v1 codes: the 3x3
median, then (unless minimal depth is set) the raw-0 hole fill, the luma-guided filter
and the temporal EMA, whose history the conditioner keeps between frames. The guidance
luma (one float in [0, 1] per pixel, the reference's guidance_luma_) is an input: None
is the all-zero guidance of a cleared conditioner.
"""
import ctypes

import numpy as np

from . import native as N
from .native import check, lib

MEDIAN, HOLES, GUIDED, EMA = 1, 2, 4, 8
ALL = MEDIAN | HOLES | GUIDED | EMA
DEFAULT_BAND = (0.3, 2.5)  # FusionHyperparams' default min_depth / max_depth


class SignalConditioner:
    def __init__(self, W=640, H=480):
        self.W, self.H = int(W), int(H)
        h = ctypes.c_void_p()
        check(lib.azu_cond_create(self.W, self.H, ctypes.byref(h)))
        self.h = h.value
        self._in = N.DeviceBuffer(self.W * self.H * 2)
        self._out = N.DeviceBuffer(self.W * self.H * 2)
        self._luma = None

    # -- state ---------------------------------------------------------------------
    def resetEMA(self, stream=None):
        check(lib.azu_cond_reset_ema(self.h, stream))

    def reset(self, stream=None):
        """reset() (SignalConditioner_omp.cpp:273-302): the depth state is the EMA history."""
        self.resetEMA(stream)

    def setMinimalDepth(self, minimal):
        check(lib.azu_cond_set_minimal_depth(self.h, 1 if minimal else 0))

    def emaState(self, stream=None):
        """emaStateForTests (SignalConditioner.h:208): the history in metres, 0 = none."""
        out = np.empty((self.H, self.W), np.float32)
        check(lib.azu_cond_read_ema(self.h, out.ctypes.data, stream))
        N.synchronize(stream)
        return out

    def setEmaState(self, state, stream=None):
        st = np.ascontiguousarray(state, np.float32)
        assert st.shape == (self.H, self.W)
        check(lib.azu_cond_write_ema(self.h, st.ctypes.data, stream))
        N.synchronize(stream)

    # -- device entry points -------------------------------------------------------
    def process_device(self, d_in, d_out, d_luma=None, band=DEFAULT_BAND, stream=None):
        check(lib.azu_cond_process(self.h, d_in, d_luma, d_out, band[0], band[1], stream))

    def run_stages_device(self, d_in, d_out, stages, d_luma=None, band=DEFAULT_BAND, stream=None):
        check(lib.azu_cond_run_stages(self.h, d_in, d_luma, d_out, band[0], band[1], int(stages), stream))

    # -- host convenience (copies in and out) ---------------------------------------
    def _host_call(self, raw, luma, fn):
        raw = np.ascontiguousarray(raw, np.uint16)
        assert raw.shape == (self.H, self.W), raw.shape
        self._in.upload(raw)
        d_luma = None
        if luma is not None:
            lu = np.ascontiguousarray(luma, np.float32)
            assert lu.shape == raw.shape
            if self._luma is None:
                self._luma = N.DeviceBuffer(self.W * self.H * 4)
            self._luma.upload(lu)
            d_luma = self._luma.ptr
        fn(self._in.ptr, self._out.ptr, d_luma)
        return self._out.download(np.uint16, raw.shape)

    def process(self, raw, luma=None, band=DEFAULT_BAND):
        """SignalConditioner::process's depth stages on a host uint16 frame; returns the
        conditioned frame (the reference conditions RawFrame::depth in place)."""
        return self._host_call(raw, luma, lambda i, o, l: self.process_device(i, o, l, band))

    def run_stages(self, raw, stages, luma=None, band=DEFAULT_BAND):
        return self._host_call(raw, luma, lambda i, o, l: self.run_stages_device(i, o, stages, l, band))

    def fillDepthHolesForTests(self, raw, band=DEFAULT_BAND):
        return self.run_stages(raw, HOLES, band=band)

    def applyDepthEmaForTests(self, raw, band=DEFAULT_BAND):
        return self.run_stages(raw, EMA, band=band)

    def close(self):
        if getattr(self, "h", None):
            lib.azu_cond_destroy(self.h)
            self.h = None
        for b in (self._in, self._out, self._luma):
            if b is not None:
                b.free()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
