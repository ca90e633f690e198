"""Batched descriptions (host side of the C ABI's nxg_batch_* entry points).

A batch is up to 12 descriptions, each with its own asset, parameters,
octaves, optional per-root attributes, outputs and workspace, generated by
one guide-preparation launch and one persistent strand launch
(nxg_generate_batch), optionally captured once into a graph and replayed
(nxg_batch_graph_*).  The reference encodes a collection as a chain of
dispatches per description (src/luisa/xgen_classic_collection.cpp:108-221);
results per description are identical to generating it alone.
"""
import ctypes as C

import numpy as np

from . import abi, check, lib, _stream_handle, _ref


def default_tuning():
    t = abi.BatchTuning()
    lib().nxg_batch_default_tuning(C.byref(t))
    return t


def tuning_candidates():
    n = lib().nxg_batch_tuning_candidates(None, 0)
    arr = (abi.BatchTuning * n)()
    lib().nxg_batch_tuning_candidates(arr, n)
    return list(arr)


class Batch:
    """Keeps every host struct an nxg_batch_item points at alive."""

    def __init__(self, device="cuda"):
        import torch
        self.torch = torch
        self.device = device
        self.entries = []  # per item: dict of the structs and tensors
        self.counter = torch.zeros(64, dtype=torch.int32, device=device)  # one word used; 256-B line of its own

    def add(self, asset, params, outputs, *, octaves=None, deformed=None, attributes=None, exprs=None, maps=None,
            workspace=None, surrogate=None, surrogate_threshold=1e-3, surrogate_stride=64):
        """attributes: float32 [N, 4] device tensor of per-root multipliers;
        with exprs (an expr.ExprSet) and maps (an expr.TexelMaps) the batch
        evaluates them into `attributes` (allocated here if None) first.
        surrogate (a surrogate.Surrogate, with exprs): the attributes are
        predicted by the network from the first texel map's features, then
        checked by the exact programs on every surrogate_stride-th root, a
        64-root tile whose check misses surrogate_threshold re-evaluated
        exactly; stats() gives (tiles re-evaluated, samples checked)."""
        torch = self.torch
        if workspace is None:
            workspace = torch.empty(max(1, asset.workspace_bytes(params)), dtype=torch.uint8, device=self.device)
        if exprs is not None and attributes is None:
            attributes = torch.empty((params.strand_count, 4), dtype=torch.float32, device=self.device)
        if surrogate is not None and exprs is None:
            raise ValueError("a surrogate needs the item's exact programs (exprs) to check it")
        e = dict(asset=asset, params=params, outputs=outputs, octaves=octaves, deformed=deformed,
                 attributes=attributes, exprs=exprs, maps=maps, workspace=workspace, out=outputs.descriptor(),
                 stencil=outputs.stencil_descriptor(), surrogate=surrogate)
        if surrogate is not None:
            tiles = (params.strand_count + 63) // 64
            e.update(surrogate_flags=torch.zeros(max(1, tiles), dtype=torch.int32, device=self.device),
                     surrogate_stats=torch.zeros(2, dtype=torch.int32, device=self.device),
                     surrogate_threshold=surrogate_threshold, surrogate_stride=surrogate_stride)
        self.entries.append(e)
        return len(self.entries) - 1

    def items(self):
        arr = (abi.BatchItem * len(self.entries))()
        for k, e in enumerate(self.entries):
            it = arr[k]
            it.asset = C.addressof(e["asset"].desc)
            it.deformed = C.addressof(e["deformed"]) if e["deformed"] is not None else None
            it.params = C.addressof(e["params"])
            it.octaves = C.addressof(e["octaves"]) if e["octaves"] is not None else None
            if e["attributes"] is not None:
                it.attributes = e["attributes"].data_ptr()
                it.attribute_capacity = e["attributes"].numel()
            if e["exprs"] is not None:
                it.exprs = C.addressof(e["exprs"].info)
                if e["maps"] is not None:
                    it.maps = C.addressof(e["maps"].structs)
                    it.map_count = e["maps"].count
            it.output = C.addressof(e["out"])
            it.stencil = C.addressof(e["stencil"]) if e["stencil"] is not None else None
            it.workspace = e["workspace"].data_ptr()
            it.workspace_bytes = e["workspace"].numel()
            if e["surrogate"] is not None:
                it.surrogate = e["surrogate"].ptr()
                it.surrogate_map = C.addressof(e["maps"].structs[0])
                it.surrogate_threshold = e["surrogate_threshold"]
                it.surrogate_stride = e["surrogate_stride"]
                it.surrogate_flags = e["surrogate_flags"].data_ptr()
                it.surrogate_flag_capacity = e["surrogate_flags"].numel()
                it.surrogate_stats = e["surrogate_stats"].data_ptr()
        self._items = arr  # keep alive while the library reads it
        return arr

    def validate(self, tuning=None):
        return lib().nxg_batch_validate(self.items(), len(self.entries), _ref(tuning))

    def generate(self, tuning=None, stream=None):
        check(lib().nxg_generate_batch(self.items(), len(self.entries), _ref(tuning),
                                       C.c_void_p(self.counter.data_ptr()), _stream_handle(stream)),
              "nxg_generate_batch")

    def graph(self, tuning=None):
        return BatchGraph(self, tuning)

    def autotune(self, candidates=None, repeats=5, stream=None):
        """Returns (best tuning, [(tuning, median ms or None if invalid)])."""
        cands = candidates if candidates is not None else tuning_candidates()
        arr = (abi.BatchTuning * len(cands))(*cands)
        times = (C.c_float * len(cands))()
        best = abi.BatchTuning()
        check(lib().nxg_batch_autotune(self.items(), len(self.entries), C.c_void_p(self.counter.data_ptr()),
                                       _stream_handle(stream), arr, len(cands), repeats, C.byref(best), times),
              "nxg_batch_autotune")
        return best, [(cands[i], None if times[i] < 0 else float(times[i])) for i in range(len(cands))]

    def stats(self, k):
        """(tiles re-evaluated, samples checked) of item k's surrogate check
        in the last run (synchronises)."""
        st = self.entries[k]["surrogate_stats"].cpu().tolist()
        return int(st[0]), int(st[1])

    @property
    def points(self):
        return sum(e["params"].strand_count * e["params"].cvs_per_strand for e in self.entries)

    @property
    def strands(self):
        return sum(e["params"].strand_count for e in self.entries)


class BatchGraph:
    """nxg_generate_batch captured once; launch() replays it on a stream."""

    def __init__(self, batch, tuning=None):
        self.batch = batch
        self.handle = C.c_void_p()
        check(lib().nxg_batch_graph_create(batch.items(), len(batch.entries), _ref(tuning),
                                           C.c_void_p(batch.counter.data_ptr()), C.byref(self.handle)),
              "nxg_batch_graph_create")

    def launch(self, stream=None):
        check(lib().nxg_batch_graph_launch(self.handle, _stream_handle(stream)), "nxg_batch_graph_launch")

    def close(self):
        if self.handle:
            check(lib().nxg_batch_graph_destroy(self.handle), "nxg_batch_graph_destroy")
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
