"""HIP-graph-captured variant chains.

A chain is a fixed sequence of plan runs on fixed device buffers -- several variants,
sizes and planes of one frame, or a prescaler followed by another pass -- captured once
into a HIP graph and replayed with a single launch.  Replays read the inputs' current
contents and overwrite the outputs, so a caller refills the inputs (or has its producer
write into them) between replays.  The library's run entry point allocates nothing and
issues no memset, so every step records as kernel nodes only.
"""
import torch


class Chain:
    """steps: iterable of (plan, x, out); the tensors must stay alive and keep their
    storage for the chain's lifetime.  ``replay(stream)`` launches the whole sequence."""

    def __init__(self, steps, device=None):
        self.steps = [(p, x, o) for p, x, o in steps]
        if not self.steps:
            raise ValueError("empty chain")
        dev = device if device is not None else self.steps[0][1].device
        side = torch.cuda.Stream(dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):  # one eager pass: first-use set-up happens outside the capture
            self._run_all()
        torch.cuda.current_stream(dev).wait_stream(side)
        torch.cuda.synchronize(dev)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._run_all()

    def _run_all(self):
        for plan, x, o in self.steps:
            plan.run(x, out=o)

    def replay(self):
        """Launch the captured sequence on the current stream."""
        self.graph.replay()

    def outputs(self):
        return [o for _, _, o in self.steps]
