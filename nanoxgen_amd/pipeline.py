"""Back-to-back generate_packed calls with the preparation overlapped.

A call of generate_packed (generate.h:612-676) is a preparation stage --
guide state, the guide offset table and the strand order
(nxg_prepare_packed_generation) -- followed by the strand kernel
(nxg_generate_packed_prepared).  The preparation is a handful of small,
latency-bound launches; the strand kernel fills every CU.  A producer of many
batches (frames, seeds, descriptions of one asset) can run call i+1's
preparation while call i's strand kernel still runs: GenerationPipeline
enqueues each preparation on a side stream into one of `depth` workspaces and
the strand kernel on the caller's stream after it.  Every call still runs
every kernel and writes exactly what generate_packed writes; only the
scheduling differs.
"""
from . import _torch, generate_packed_prepared, prepare_packed


class GenerationPipeline:
    """Enqueues generate_packed calls of one asset with each call's
    preparation on a side stream (default priority) into workspace
    i % depth, so that it overlaps the previous call's strand kernel.

    `params` sizes the workspaces: every later call's strand count and CV
    count must need no more workspace than it.  `stream` (torch stream,
    default the current one) receives the strand kernels; outputs written by
    submit() are ready on it, in submission order.  Work of the caller on
    other streams that reads the outputs must wait on `stream` as usual.
    A preparation reads the asset, the deformed geometry and the parameters
    after the caller's stream as it stood at construction, and after the
    `after` event of its submit(): work that rewrites the deformed geometry
    between calls passes its event there."""

    def __init__(self, asset, params, *, octaves=None, deformed=None, depth=2, stream=None, attribute_fn=None,
                 attributes_after_prepare=False):
        torch = _torch()
        if depth < 1:
            raise ValueError("GenerationPipeline: depth must be at least 1")
        self.asset, self.params, self.octaves, self.deformed = asset, params, octaves, deformed
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        self.side = torch.cuda.Stream(device=asset.tensor.device)
        self.capacity = max(1, asset.workspace_bytes(params))
        self.workspaces = [torch.empty(self.capacity, dtype=torch.uint8, device=asset.tensor.device)
                           for _ in range(depth)]
        # per-root attribute multipliers (nxg_generate_packed_attributed):
        # attribute_fn(params, attributes, stream) fills slot k's float32
        # [N, 4] buffer on the side stream before the preparation, or with
        # attributes_after_prepare attribute_fn(params, attributes, stream,
        # workspace) after it (a producer that reads the prepared strand
        # order, such as the surrogate)
        self.attribute_fn = attribute_fn
        self.attributes_after_prepare = attributes_after_prepare
        self.attributes = ([torch.empty((params.strand_count, 4), dtype=torch.float32, device=asset.tensor.device)
                            for _ in range(depth)] if attribute_fn is not None else None)
        self.ready = [torch.cuda.Event() for _ in range(depth)]
        self.done = [torch.cuda.Event() for _ in range(depth)]  # after the last strand kernel reading slot k
        self.calls = 0
        # the side stream starts after whatever the caller's stream holds
        self.side.wait_stream(self.stream)

    def submit(self, outputs, params=None, *, after=None, events=None):
        """Enqueue one generate_packed call writing `outputs` (PackedOutputs).
        `params` (default the constructor's) may change the seed and the
        other generation parameters.  With an attribute_fn, the call's
        attributes are evaluated into slot i % depth on the side stream
        before its preparation.  `after` (optional torch event): work the
        preparation must follow.  `events` (optional): four torch events
        recorded around the preparation (side stream) and the strand kernel
        (caller's stream)."""
        params = params if params is not None else self.params
        need = self.asset.workspace_bytes(params)
        if need > self.capacity:
            raise ValueError(f"GenerationPipeline: this call needs {need} workspace bytes, "
                             f"the pipeline holds {self.capacity} per slot")
        k = self.calls % len(self.workspaces)
        ws = self.workspaces[k]
        if self.calls >= len(self.workspaces):
            self.side.wait_event(self.done[k])
        if after is not None:
            self.side.wait_event(after)
        if events:
            events[0].record(self.side)
        attributes = None
        if self.attribute_fn is not None:
            if params.strand_count > self.attributes[k].shape[0]:
                raise ValueError(f"GenerationPipeline: {params.strand_count} strands, the attribute slots hold "
                                 f"{self.attributes[k].shape[0]}")
            attributes = self.attributes[k][:params.strand_count]
            if not self.attributes_after_prepare:
                self.attribute_fn(params, attributes, self.side)
        prepare_packed(self.asset, params, ws, deformed=self.deformed, stream=self.side)
        if attributes is not None and self.attributes_after_prepare:
            self.attribute_fn(params, attributes, self.side, ws)
        if events:
            events[1].record(self.side)
        self.ready[k].record(self.side)
        self.stream.wait_event(self.ready[k])
        if events:
            events[2].record(self.stream)
        generate_packed_prepared(self.asset, params, outputs, ws, octaves=self.octaves, deformed=self.deformed,
                                 stream=self.stream, attributes=attributes)
        if events:
            events[3].record(self.stream)
        self.done[k].record(self.stream)
        self.calls += 1
        return outputs

    def drain(self):
        """Make the caller's stream wait for every enqueued preparation (the
        strand kernels already order after theirs)."""
        self.stream.wait_stream(self.side)
