"""Per-kernel timing with HIP events recorded on the launch stream.

The rasterizer's kernels run on torch's current stream, so events recorded on
that stream before and after an entry point bracket exactly that entry
point's kernels.  Disabled by default (no overhead)."""
import contextlib

import torch


class Spans:
    """enabled turns recording on; only (a set of names, or None for all)
    limits it to some spans.  Each recorded event is a barrier packet on the
    stream, so a timed region should record as few as it needs."""

    def __init__(self):
        self.enabled = False
        self.only = None
        self.events = {}

    def reset(self):
        self.events = {}

    @contextlib.contextmanager
    def span(self, name):
        if not self.enabled or (self.only is not None and name not in self.only):
            yield
            return
        s = torch.cuda.current_stream()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(s)
        try:
            yield
        finally:
            b.record(s)
            self.events.setdefault(name, []).append((a, b))

    def summary(self):
        """{name: (launches, total_ms, mean_ms)} (synchronises)."""
        torch.cuda.synchronize()
        out = {}
        for k, v in self.events.items():
            ms = [a.elapsed_time(b) for a, b in v]
            out[k] = (len(ms), sum(ms), sum(ms) / max(1, len(ms)))
        return out


SPANS = Spans()
