"""Batched launch: many forward problems (phantoms of uneven slice counts, angle counts and energy bins) in one
kernel launch per voxel type (a flat grid or a persistent grid with per-XCD work queues, chosen by the
autotuner), optionally followed by per-problem FDK plans, capturable in a hipGraph."""
from __future__ import annotations

import ctypes as C

from ._lib import check, lib
from ._lib import stream_arg as _stream


class Batch:
    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().vcbct_batch_create(C.byref(h), int(device)))
        self.h = h
        self._keep = []

    def add_forward(self, projector, out, xfactor):
        """projector: a Projector with its volume and spectrum set; out: FP32 device tensor of its projections."""
        check(lib().vcbct_batch_add_forward(self.h, projector.h, C.c_void_p(out.data_ptr()), float(xfactor)))
        self._keep += [projector, out]

    def add_fdk(self, plan, proj, filtered, recon):
        plan._check(proj, plan.shape_proj, "proj")
        plan._check(filtered, plan.shape_filtered, "filtered")
        plan._check(recon, plan.shape_recon, "recon")
        check(lib().vcbct_batch_add_fdk(self.h, plan.h, C.c_void_p(proj.data_ptr()), C.c_void_p(filtered.data_ptr()),
                                        C.c_void_p(recon.data_ptr()), recon.numel() * 4))
        self._keep += [plan, proj, filtered, recon]

    def run(self, stream=None):
        check(lib().vcbct_batch_run(self.h, _stream(stream)))

    def capture(self):
        check(lib().vcbct_batch_capture_graph(self.h))

    def replay(self, stream=None):
        check(lib().vcbct_batch_launch_graph(self.h, _stream(stream)))

    @property
    def blocks(self):
        """Grid in use: 0 = flat, n > 0 = persistent with n blocks, -2 = sequence of plain per-problem launches
        (after the first run)."""
        return lib().vcbct_batch_blocks(self.h)

    def set_grid(self, blocks):
        """Fix the grid instead of autotuning: 0 = flat, n > 0 = persistent with n blocks, -2 = sequence of plain
        per-problem launches (each with its geometry's tuned tile shape), -1 = autotune."""
        check(lib().vcbct_batch_set_grid(self.h, int(blocks)))

    def set_tile(self, row_groups=0, cols=0, order=0, split=0):
        """Fix the tile shape instead of autotuning it: the flat grid's 4-wave workgroup (row_groups x cols = 4)
        and order, and for both grids the split (0: columns, 1: angles divided among the XCDs); row_groups=0
        hands it back to the autotuner."""
        check(lib().vcbct_batch_set_tile(self.h, int(row_groups), int(cols), int(order), int(split)))

    @property
    def tile(self):
        """(row_groups, cols, order, split) in use (after the first run)."""
        a = (C.c_int * 4)()
        check(lib().vcbct_batch_get_tile(self.h, a))
        return tuple(a)

    def close(self):
        if getattr(self, "h", None):
            lib().vcbct_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune_log():
    buf = C.create_string_buffer(1 << 16)
    lib().vcbct_autotune_log(buf, len(buf))
    return buf.value.decode()
