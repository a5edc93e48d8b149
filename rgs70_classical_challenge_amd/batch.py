"""Many uneven graph-state instances in one captured HIP graph.

Each instance (n qubits, its graph and unitaries) gets a state slab, mass
pyramid, uniforms and outputs.  Instances are packed into memory waves whose
total device footprint stays under a budget.  ``run_single`` places every
wave in one shared arena (wave k+1 reuses wave k's memory) and captures the
whole batch as ONE graph: each wave forks its instances over several streams
(default priority) and joins them before the next wave starts, so the batch
is a single ``graphLaunch`` from the host.  ``run_graph`` captures one graph
per wave; ``run_eager`` is the one-launch-sequence-per-instance baseline.
``run_fused`` hands each wave to the native multi-state batch
(rgs_slab_batch_*): instances with the same pass schedule share every pass's
launch (k_tile_multi), and all pyramids and draws of a wave take a few
launches (k_pyramid_multi, k_draw_multi), so the launch count no longer grows
with the number of instances.
Tile shapes per n come from the autotuner table (``autotune.py``) when
present.  The reference runs one state per process invocation
(work/rgs_attack/born_production.py:44-58); no batched counterpart exists.
"""
import numpy as np

from . import _lib
from ._lib import call, ptr
from .engine import GraphStateSampler
from .instances import instance


def uneven_instances(count=256, seed=20261204, nmin=16, nmax=30):
    """n ~ U{nmin..nmax} per instance (seeded); 3-regular graphs for even n,
    Erdos-Renyi G(n, 3/n) for odd n; QR unitaries."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ns = rng.integers(nmin, nmax + 1, count)
    seeds = rng.integers(0, 2 ** 62, count)
    return [(int(n), int(s)) for n, s in zip(ns, seeds)]


def footprint(n, samples, mass_bits=8):
    return (8 << n) + (16 << max(0, n - mass_bits)) + samples * 24 + (1 << 16)


class Arena:
    """One device allocation handed out as aligned views (take), rewound
    between memory waves (reset)."""

    ALIGN = 256

    def __init__(self, nbytes, device=0):
        import torch
        self.buf = torch.empty(nbytes, dtype=torch.uint8, device=torch.device('cuda', device))
        self.offset = 0
        self.high = 0

    def take(self, count, dtype):
        import torch
        size = count * torch.empty((), dtype=dtype).element_size()
        start = -(-self.offset // self.ALIGN) * self.ALIGN
        if start + size > self.buf.numel():
            raise MemoryError('arena exhausted')
        self.offset = start + size
        self.high = max(self.high, self.offset)
        return self.buf[start:start + size].view(dtype)

    def reset(self):
        self.offset = 0


class BatchSampler:
    def __init__(self, specs, samples=1000, budget_bytes=200 << 30, streams=4, tuning='auto', seed=1):
        import torch
        self.torch = torch
        self.specs = list(specs)
        self.samples = samples
        self.streams = [torch.cuda.Stream() for _ in range(streams)]
        self.tuning = tuning
        self.rng = np.random.Generator(np.random.PCG64(seed))
        order = sorted(range(len(self.specs)), key=lambda i: -self.specs[i][0])
        self.waves = []
        cur, used = [], 0
        for i in order:
            need = footprint(self.specs[i][0], samples)
            if cur and used + need > budget_bytes:
                self.waves.append(cur)
                cur, used = [], 0
            cur.append(i)
            used += need
        if cur:
            self.waves.append(cur)
        self.graphs = {}

    def _tuning_for(self, n):
        # 'auto': the autotuner's table for this device (default_tuning where it has no entry);
        # 'default': the fixed default_tuning for every size, the untuned baseline
        if self.tuning == 'auto':
            return None
        if self.tuning == 'default':
            from .engine import default_tuning
            return default_tuning(n)
        return self.tuning

    def _build(self, wave, arena=None):
        torch = self.torch
        items = []
        for i in wave:
            n, seed = self.specs[i]
            g, u = instance(n, seed)
            s = GraphStateSampler(g, u, pipeline='slab', tuning=self._tuning_for(n), arena=arena)
            uni = torch.as_tensor(self.rng.random(self.samples)).to(s.slab.device)
            idx = torch.empty(self.samples, dtype=torch.int64, device=s.slab.device)
            pr = torch.empty(self.samples, dtype=torch.float64, device=s.slab.device)
            items.append(dict(i=i, s=s, uni=uni, idx=idx, p=pr))
        return items

    def _launch(self, item, stream):
        s = item['s']
        sl = s.slab
        h = stream.cuda_stream
        from .engine import ctypes_ref
        call('rgs_slab_pipeline', ctypes_ref(s._desc), h)
        call('rgs_mass_pyramid', ptr(sl.pyramid), sl.base_bits, h)
        call('rgs_draw', ptr(sl.state), s.n, ptr(sl.pyramid), sl.mass_bits, ptr(item['uni']), self.samples,
             s.flip, ptr(item['idx']), ptr(item['p']), h)

    def run_eager(self):
        """One launch sequence per instance on the current stream, wave by wave."""
        torch = self.torch
        out = {}
        for wave in self.waves:
            items = self._build(wave)
            cur = torch.cuda.current_stream()
            for it in items:
                self._launch(it, cur)
            torch.cuda.synchronize()
            for it in items:
                out[it['i']] = it['idx'].cpu().numpy()
            del items
            torch.cuda.empty_cache()
        return out

    def _capture(self, items):
        """One graph for a wave: fork onto the side streams, launch every
        instance, join back."""
        torch = self.torch
        main = torch.cuda.Stream()
        graph = torch.cuda.CUDAGraph()
        # warm the launch path once outside capture (one-time attribute setup)
        for k, it in enumerate(items):
            self._launch(it, self.streams[k % len(self.streams)])
        torch.cuda.synchronize()
        with torch.cuda.stream(main):
            graph.capture_begin()
            for st in self.streams:
                st.wait_stream(main)
            for k, it in enumerate(items):
                self._launch(it, self.streams[k % len(self.streams)])
            for st in self.streams:
                main.wait_stream(st)
            graph.capture_end()
        return graph

    def run_graph(self, replays=1, keep=False):
        """Capture each wave as one multi-stream graph and replay it."""
        torch = self.torch
        out = {}
        for wi, wave in enumerate(self.waves):
            items = self._build(wave)
            graph = self._capture(items)
            for _ in range(replays):
                graph.replay()
            torch.cuda.synchronize()
            for it in items:
                out[it['i']] = it['idx'].cpu().numpy()
            if keep:
                self.graphs[wi] = (graph, items)
            else:
                del graph, items
                torch.cuda.empty_cache()
        return out

    def _capture_all(self, waves):
        """One graph for the whole batch: each wave forks onto the side
        streams and joins back before the next wave (which reuses its memory)
        is launched."""
        torch = self.torch
        main = torch.cuda.Stream()
        graph = torch.cuda.CUDAGraph()
        for items in waves:                  # warm the launch path outside capture
            for k, it in enumerate(items):
                self._launch(it, self.streams[k % len(self.streams)])
            torch.cuda.synchronize()
        with torch.cuda.stream(main):
            graph.capture_begin()
            for items in waves:
                for st in self.streams:
                    st.wait_stream(main)
                for k, it in enumerate(items):
                    self._launch(it, self.streams[k % len(self.streams)])
                for st in self.streams:
                    main.wait_stream(st)
            graph.capture_end()
        return graph

    def build_single(self):
        """Every wave's instances with their states in one shared arena (sized
        for the largest wave), and the batch captured as one graph."""
        sizes = [sum(footprint(self.specs[i][0], self.samples) for i in wave) for wave in self.waves]
        self.arena = Arena(max(sizes) + (64 << 20))
        waves = []
        for wave in self.waves:
            self.arena.reset()
            waves.append(self._build(wave, arena=self.arena))
        self.single = (self._capture_all(waves), waves)
        return self.single

    def run_single(self, replays=1):
        """Replay the one-graph batch; returns {instance: drawn indices}."""
        graph, waves = self.single if getattr(self, 'single', None) else self.build_single()
        for _ in range(replays):
            graph.replay()
        self.torch.cuda.synchronize()
        return {it['i']: it['idx'].cpu().numpy() for items in waves for it in items}

    def build_fused(self):
        """Every wave as one multi-state batch (rgs_slab_batch_create): the
        wave's instances with the same pass schedule (same n and tile plan)
        share each pass's launch, and the pyramids and draws of the whole
        wave take a few launches more.  Waves share one arena, as in
        build_single; each wave's kernel arguments are uploaded once here."""
        import ctypes
        self.close_fused()
        sizes = [sum(footprint(self.specs[i][0], self.samples) for i in wave) for wave in self.waves]
        self.arena = Arena(max(sizes) + (64 << 20))
        self.fused = []
        for wave in self.waves:
            self.arena.reset()
            items = self._build(wave, arena=self.arena)
            arr = (_lib.SlabBatchItem * len(items))()
            for k, it in enumerate(items):
                arr[k].desc = ctypes.pointer(it['s']._desc)
                arr[k].uniforms = ptr(it['uni'])
                arr[k].samples = self.samples
                arr[k].out_index = ptr(it['idx'])
                arr[k].out_p = ptr(it['p'])
                arr[k].flip = it['s'].flip
            h = ctypes.c_void_p()
            call('rgs_slab_batch_create', arr, len(items), ctypes.byref(h))
            self.fused.append((h, items))
        return self.fused

    def fused_launches(self):
        """(kernel launches, pass-schedule groups) summed over the waves."""
        import ctypes
        launches = groups = 0
        for h, _ in self.fused:
            a, b = ctypes.c_int(), ctypes.c_int()
            call('rgs_slab_batch_info', h, ctypes.byref(a), ctypes.byref(b))
            launches += a.value
            groups += b.value
        return launches, groups

    def launch_fused(self, stream=None):
        """Enqueue every wave's batch on one stream (the current one by
        default); waves run in order, so a wave may reuse the arena."""
        st = (stream or self.torch.cuda.current_stream()).cuda_stream
        for h, _ in self.fused:
            call('rgs_slab_batch_run', h, st)

    def run_fused(self, replays=1):
        """Run the fused batch; returns {instance: drawn indices}."""
        if not getattr(self, 'fused', None):
            self.build_fused()
        for _ in range(replays):
            self.launch_fused()
        self.torch.cuda.synchronize()
        return {it['i']: it['idx'].cpu().numpy() for _, items in self.fused for it in items}

    def capture_fused(self):
        """The fused batch captured as one graph (its launches only)."""
        torch = self.torch
        if not getattr(self, 'fused', None):
            self.build_fused()
        self.launch_fused()                 # warm outside capture
        torch.cuda.synchronize()
        main = torch.cuda.Stream()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.stream(main):
            graph.capture_begin()
            self.launch_fused(main)
            graph.capture_end()
        return graph

    def close_fused(self):
        for h, _ in getattr(self, 'fused', None) or []:
            call('rgs_slab_batch_destroy', h)
        self.fused = None

    def benchmark(self, reps=3):
        """Device time of the whole batch (state, masses, draws of every
        instance): one launch sequence per instance on one stream, against
        one graph replay per wave.  Building the instances (allocation and
        host tables) is outside both timings."""
        import time
        torch = self.torch
        eager_s = graph_s = 0.0
        for wave in self.waves:
            items = self._build(wave)
            cur = torch.cuda.current_stream()
            for it in items:
                self._launch(it, cur)
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(reps):
                for it in items:
                    self._launch(it, cur)
            torch.cuda.synchronize()
            eager_s += (time.perf_counter() - t) / reps
            graph = self._capture(items)
            graph.replay()
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(reps):
                graph.replay()
            torch.cuda.synchronize()
            graph_s += (time.perf_counter() - t) / reps
            del graph, items
            torch.cuda.empty_cache()
        # the whole batch as one graph launch (waves share one arena)
        graph, _ = self.build_single()
        graph.replay()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            graph.replay()
        torch.cuda.synchronize()
        single_s = (time.perf_counter() - t) / reps
        self.single = None
        del graph
        self.arena = None
        torch.cuda.empty_cache()
        # the fused batch: instances sharing a pass schedule share launches
        self.build_fused()
        launches, groups = self.fused_launches()
        self.launch_fused()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            self.launch_fused()
        torch.cuda.synchronize()
        fused_s = (time.perf_counter() - t) / reps
        fgraph = self.capture_fused()
        fgraph.replay()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fgraph.replay()
        torch.cuda.synchronize()
        fused_graph_s = (time.perf_counter() - t) / reps
        del fgraph
        self.close_fused()
        self.arena = None
        torch.cuda.empty_cache()
        total = len(self.specs) * self.samples
        amps = sum(1 << n for n, _ in self.specs)
        return {'instances': len(self.specs), 'samples_per_instance': self.samples, 'waves': len(self.waves),
                'streams': len(self.streams), 'amplitudes': amps,
                'eager_ms': round(eager_s * 1e3, 3), 'graph_ms': round(graph_s * 1e3, 3),
                'single_graph_ms': round(single_s * 1e3, 3),
                'eager_samples_per_s': round(total / eager_s, 1), 'graph_samples_per_s': round(total / graph_s, 1),
                'single_graph_samples_per_s': round(total / single_s, 1),
                'graph_speedup': round(eager_s / graph_s, 3), 'single_graph_speedup': round(eager_s / single_s, 3),
                'fused_ms': round(fused_s * 1e3, 3), 'fused_graph_ms': round(fused_graph_s * 1e3, 3),
                'fused_launches': launches, 'fused_groups': groups,
                'fused_samples_per_s': round(total / fused_s, 1),
                'fused_speedup': round(eager_s / fused_s, 3), 'fused_graph_speedup': round(eager_s / fused_graph_s, 3)}


def main(argv):
    """python -m rgs70_classical_challenge_amd.batch [count] [nmin] [nmax]: one JSON line."""
    import json
    count = int(argv[0]) if argv else 256
    nmin = int(argv[1]) if len(argv) > 1 else 16
    nmax = int(argv[2]) if len(argv) > 2 else 30
    specs = uneven_instances(count, nmin=nmin, nmax=nmax)
    res = BatchSampler(specs, samples=1000).benchmark()
    res['workload'] = f'{count} instances, n uniform in {nmin}..{nmax}, 1000 samples each'
    # the same batch with the fixed default tile choices: what the autotuner's table buys
    base = BatchSampler(specs, samples=1000, tuning='default').benchmark(reps=1)
    res['default_tuning'] = {k: base[k] for k in ('eager_ms', 'graph_ms', 'single_graph_ms')}
    res['autotuner_speedup'] = round(base['single_graph_ms'] / res['single_graph_ms'], 3)
    # the launch-bound part of the batch alone (its small instances), where the graph matters
    small = [sp for sp in specs if sp[0] <= 22]
    if small:
        sres = BatchSampler(small, samples=1000).benchmark()
        res['small_instances'] = {'n_max': 22, 'instances': len(small),
                                  **{k: sres[k] for k in ('eager_ms', 'graph_ms', 'single_graph_ms', 'graph_speedup',
                                                          'single_graph_speedup', 'fused_ms', 'fused_graph_ms',
                                                          'fused_launches', 'fused_groups', 'fused_speedup',
                                                          'fused_graph_speedup')}}
    print(json.dumps(res), flush=True)


if __name__ == '__main__':
    import sys
    main(sys.argv[1:])
