"""A whole slice-splat training step over a batch of frames as one captured hipGraph.

Per step (all on one stream, no host synchronisation, no allocation):
  us3d_ug_map (its pack applies the opacity/intensity sigmoids, pipeline_god_file.py:369-373) ->
  us3d_ug_forward -> us3d_image_loss (L1 + w (1 - global SSIM) and its gradient, :606-637) ->
  us3d_ug_backward (its reduce returns the gradients w.r.t. the logits) -> us3d_adam_step for each parameter
  group (torch.optim.Adam semantics, :740,764) -> us3d_adam_advance.
The instance arrays have a fixed capacity (from an eager mapping of the same batch, with headroom). If a step's
mapping overflows it, the Adam kernels skip the update (they read the overflow flag), `step()` notices it when it
checks the stats, grows the capacity, captures again and repeats the step -- so an overflow never corrupts the
parameters. The overflow check synchronises; `check_every` sets how often it runs.
"""

from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence, Tuple

import torch

from . import _lib
from .batch import I32_LIMIT, CapacityHint, FrameSpec, SweepSpec, new_workspace, read_stats
from .ultragauss import UGBatch, UltraGaussRenderConfig


def frame_chunks(frames: Sequence[FrameSpec], sweeps: Sequence[SweepSpec], chunk_frames: Optional[int]):
    """Consecutive runs of at most chunk_frames frames (all frames when None), each with its sweeps renumbered
    compactly: [(frame begin, frame end, frames, sweeps)]."""
    from dataclasses import replace

    n = len(frames)
    step = n if not chunk_frames else max(1, int(chunk_frames))
    out = []
    for a in range(0, n, step):
        b = min(n, a + step)
        used: List[int] = []
        sub = []
        for f in frames[a:b]:
            if f.sweep not in used:
                used.append(f.sweep)
            sub.append(replace(f, sweep=used.index(f.sweep)))
        out.append((a, b, sub, [sweeps[s] for s in used]))
    return out


class SliceSplatTrainer:
    """The captured training step. With chunk_frames, the frames are rendered chunk by chunk inside the same step
    (map -> forward -> loss -> adjoint per chunk of at most chunk_frames frames, the chunks' gradients summed, then
    one Adam update), so the workspace holds one chunk's instances and pairs: memory is bounded by the chunk, not
    the sweep. The loss is the mean over all frames either way, so the update is the same.
    fused_loss: the loss statistics are summed inside the forward splat and the loss gradient is formed inside the
    adjoint (us3d_ug_forward_loss / us3d_ug_backward_loss); False runs the separate loss kernels
    (us3d_image_loss) between them. Both give the same step up to FP64 summation order."""

    def __init__(self, means: torch.Tensor, tril: torch.Tensor, raw_opacity: torch.Tensor,
                 raw_intensity: torch.Tensor, frames: Sequence[FrameSpec], target: torch.Tensor,
                 cfg: UltraGaussRenderConfig, sweeps: Optional[Sequence[SweepSpec]] = None, lr: float = 1e-3,
                 betas=(0.9, 0.999), eps: float = 1e-8, w_l1: float = 1.0, w_ssim: float = 0.2,
                 headroom: float = 1.3, use_graph: bool = True, check_every: int = 1,
                 chunk_frames: Optional[int] = None, fused_loss: bool = True):
        dev = means.device
        self.fused_loss = bool(fused_loss)
        self.dev = dev
        G = means.shape[0]
        f32 = dict(dtype=torch.float32, device=dev)
        self.means = means.detach().to(**f32).contiguous().clone()
        self.tril = tril.detach().to(**f32).contiguous().clone()
        self.raw_o = raw_opacity.detach().to(**f32).reshape(G).contiguous().clone()
        self.raw_c = raw_intensity.detach().to(**f32).reshape(G).contiguous().clone()
        self.params = [self.means, self.tril, self.raw_o, self.raw_c]
        self.exp_avg = [torch.zeros_like(p) for p in self.params]
        self.exp_avg_sq = [torch.zeros_like(p) for p in self.params]
        self.step_count = torch.zeros(1, dtype=torch.int32, device=dev)
        self.lr, self.betas, self.eps = float(lr), betas, float(eps)
        self.w_l1, self.w_ssim = float(w_l1), float(w_ssim)
        self.cfg = cfg
        settings = cfg.settings()
        settings.raw_activations = 1  # opacity and intensity are passed as logits
        sweeps = list(sweeps or [SweepSpec(0, G)])
        self.F = len(frames)
        self.chunks = []  # (frame begin, frame end, pixel begin, pixel end, UGBatch)
        px0 = 0
        for a, b, sub, sub_sweeps in frame_chunks(frames, sweeps, chunk_frames):
            batch = UGBatch(sub, sub_sweeps, G, settings, dev, CapacityHint())
            n = batch.layout.n_pixels
            self.chunks.append((a, b, px0, px0 + n, batch))
            px0 += n
        self.batch = self.chunks[0][4]
        P = px0
        Pc = max(c[3] - c[2] for c in self.chunks)
        self.target = target.detach().to(**f32).reshape(P).contiguous()
        self.image = torch.empty(Pc, **f32)
        self.alpha = torch.empty(Pc, **f32)
        self.aux = torch.empty((Pc, 4), **f32)
        self.d_image = torch.empty(Pc, **f32)
        self.loss = torch.empty(self.F, **f32)
        fc = max(c[1] - c[0] for c in self.chunks)
        self.loss_tmp = torch.empty(int(_lib.lib().us3d_image_loss_tmp_bytes(fc)) // 8, dtype=torch.float64,
                                    device=dev)
        self.grads = [torch.empty_like(p) for p in self.params]
        self.grads_chunk = [torch.empty_like(p) for p in self.params] if len(self.chunks) > 1 else self.grads
        self.skip = torch.zeros(1, dtype=torch.int32, device=dev)  # any chunk of the step overflowed
        self.headroom = headroom
        self.use_graph = use_graph
        self.check_every = max(1, int(check_every))
        self.n_steps = 0
        self.graph = None
        self._allocate(self._sizing_capacity())

    def _sizing_capacity(self) -> Tuple[int, int]:
        # size the instance and visible-pair arrays from eager mappings of the current parameters (largest chunk),
        # releasing each mapping's workspace before the next
        k = v = 0
        with torch.no_grad():
            for c in self.chunks:
                st = c[4].map(self.means, self.tril, self.raw_o, self.raw_c).stats
                k, v = max(k, int(st.n_instances)), max(v, int(st.n_visible))
        return int(k * self.headroom) + 4096, int(v * self.headroom) + 1024

    # ------------------------------------------------------------------ plumbing
    def _allocate(self, caps: Tuple[int, int]) -> None:
        self.k_capacity = min(int(caps[0]), I32_LIMIT)
        self.v_capacity = max(1, min(int(caps[1]), max(c[4].layout.n_pairs for c in self.chunks), I32_LIMIT))
        self.graph = None
        self.workspace = None  # the old workspace goes before the new one is allocated
        self.workspace = new_workspace(max(c[4].workspace_bytes(self.k_capacity, self.v_capacity)
                                           for c in self.chunks), self.dev)
        self.plans = [c[4].tables.plan(self.workspace, self.k_capacity, self.v_capacity) for c in self.chunks]
        fuse_bytes = max(int(_lib.lib().us3d_ug_loss_tmp_bytes(ctypes.byref(p))) for p in self.plans)
        self.fuse_tmp = torch.empty(fuse_bytes // 8, dtype=torch.float64, device=self.dev)
        off = self.batch.layout_offsets(self.k_capacity, self.v_capacity)
        self.overflow_ptr = self.workspace.data_ptr() + off.stats + _lib.Stats.overflow.offset
        self.overflow_view = self.workspace[off.stats + _lib.Stats.overflow.offset:
                                            off.stats + _lib.Stats.overflow.offset + 4].view(torch.int32)

    def _body(self) -> None:
        L = _lib.lib()
        s = _lib.stream_ptr(self.dev)
        chunked = len(self.chunks) > 1
        if chunked:
            self.skip.zero_()
        for ci, ((a, b, p0, p1, batch), plan) in enumerate(zip(self.chunks, self.plans)):
            st = batch.settings
            nf = b - a
            _lib.check(L.us3d_ug_map(plan, _lib.ptr(self.means), _lib.ptr(self.tril), _lib.ptr(self.raw_o),
                                     _lib.ptr(self.raw_c), st, s), "us3d_ug_map")
            if chunked:
                torch.maximum(self.skip, self.overflow_view, out=self.skip)
            out = self.grads if ci == 0 else self.grads_chunk
            gt = _lib.ptr(self.target) + 4 * p0
            if self.fused_loss:
                _lib.check(L.us3d_ug_forward_loss(plan, None, st, _lib.ptr(self.image), _lib.ptr(self.alpha), None,
                                                  _lib.ptr(self.aux), gt, self.w_l1, self.w_ssim, 1.0 / self.F,
                                                  _lib.ptr(self.loss) + 4 * a, _lib.ptr(self.fuse_tmp), s),
                           "us3d_ug_forward_loss")
                _lib.check(L.us3d_ug_backward_loss(plan, None, _lib.ptr(self.tril), st, _lib.ptr(self.aux), gt,
                                                   _lib.ptr(self.fuse_tmp), _lib.ptr(out[0]), _lib.ptr(out[1]),
                                                   _lib.ptr(out[2]), _lib.ptr(out[3]), s), "us3d_ug_backward_loss")
            else:
                _lib.check(L.us3d_ug_forward(plan, None, st, _lib.ptr(self.image), _lib.ptr(self.alpha),
                                             None, _lib.ptr(self.aux), s), "us3d_ug_forward")
                _lib.check(L.us3d_image_loss(_lib.ptr(self.image), gt, _lib.ptr(batch.tables.pixel_base), nf,
                                             self.w_l1, self.w_ssim, 1.0 / self.F, _lib.ptr(self.loss) + 4 * a,
                                             _lib.ptr(self.d_image), _lib.ptr(self.loss_tmp), s), "us3d_image_loss")
                _lib.check(L.us3d_ug_backward(plan, None, _lib.ptr(self.tril), st, _lib.ptr(self.aux),
                                              _lib.ptr(self.d_image), None, None, _lib.ptr(out[0]), _lib.ptr(out[1]),
                                              _lib.ptr(out[2]), _lib.ptr(out[3]), s), "us3d_ug_backward")
            if ci > 0:
                for g, gc in zip(self.grads, self.grads_chunk):
                    g.add_(gc)
        skip = self.skip.data_ptr() if chunked else self.overflow_ptr
        b1, b2 = self.betas
        for p, g, m, v in zip(self.params, self.grads, self.exp_avg, self.exp_avg_sq):
            _lib.check(L.us3d_adam_step(_lib.ptr(p), _lib.ptr(g), _lib.ptr(m), _lib.ptr(v), p.numel(),
                                        _lib.ptr(self.step_count), skip, self.lr, b1, b2, self.eps, s),
                       "us3d_adam_step")
        _lib.check(L.us3d_adam_advance(_lib.ptr(self.step_count), skip, s), "us3d_adam_advance")

    def _capture(self) -> None:
        side = torch.cuda.Stream(device=self.dev)
        side.wait_stream(torch.cuda.current_stream(self.dev))
        with torch.cuda.stream(side):
            pass  # no eager warm-up inside the graph step: it would apply an Adam update
        torch.cuda.current_stream(self.dev).wait_stream(side)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._body()
        self.graph = g

    def overflowed(self) -> bool:
        """True if any step since the workspace was allocated (or since clear_overflow) overflowed: the sticky
        us3d_stats.overflow_any flag, which every overflowing map sets and none clears (synchronises)."""
        return bool(read_stats(self.workspace).overflow_any)

    def clear_overflow(self) -> None:
        self.workspace[: ctypes.sizeof(_lib.Stats)].zero_()

    # ------------------------------------------------------------------ public
    def step(self) -> None:
        """One training step; repeats the step with a larger capacity if the mapping overflowed."""
        while True:
            if self.use_graph:
                if self.graph is None:
                    self._capture()
                self.graph.replay()
            else:
                self._body()
            self.n_steps += 1
            if self.n_steps % self.check_every != 0:
                return
            if len(self.chunks) == 1:
                st = read_stats(self.workspace)
                if not st.overflow:
                    return
                k = int(st.n_instances * self.headroom) + 4096 if st.n_instances > self.k_capacity else self.k_capacity
                v = int(st.n_visible * self.headroom) + 1024 if st.n_visible > self.v_capacity else self.v_capacity
                self._allocate((k, v))
            else:
                if not int(self.skip.item()):
                    return
                self._allocate(self._sizing_capacity())

    def state(self) -> List[torch.Tensor]:
        return [p.detach() for p in self.params]

    def losses(self) -> torch.Tensor:
        return self.loss.detach().clone()


class GaussianPileTrainer:
    """The same captured step for the GaussianPile us_focus splat (method "gaussianpile",
    pipeline_god_file.py:435-452): us3d_gp_map with raw activations (its pack applies the opacity sigmoid and
    the scales' max(exp(x), 1e-5), :360-370; the kernels normalise the quaternions as :367 does) ->
    us3d_gp_forward -> us3d_image_loss -> us3d_gp_backward (gradients w.r.t. the raw values) -> us3d_adam_step
    per group -> us3d_adam_advance. Intensities are one, as the facade passes them. Overflow handling,
    `check_every` and `fused_loss` (us3d_gp_forward_loss / us3d_gp_backward_loss in place of the three calls) as
    in SliceSplatTrainer."""

    def __init__(self, means: torch.Tensor, raw_scales: torch.Tensor, raw_quats: torch.Tensor,
                 raw_opacity: torch.Tensor, frames: Sequence[FrameSpec], target: torch.Tensor, cfg,
                 sweeps: Optional[Sequence[SweepSpec]] = None, lr: float = 1e-3, betas=(0.9, 0.999),
                 eps: float = 1e-8, w_l1: float = 1.0, w_ssim: float = 0.2, headroom: float = 1.3,
                 use_graph: bool = True, check_every: int = 1, fused_loss: bool = True):
        from .gaussianpile import GPBatch

        self.fused_loss = bool(fused_loss)
        dev = means.device
        self.dev = dev
        G = means.shape[0]
        f32 = dict(dtype=torch.float32, device=dev)
        self.means = means.detach().to(**f32).contiguous().clone()
        self.raw_s = raw_scales.detach().to(**f32).contiguous().clone()
        self.raw_q = raw_quats.detach().to(**f32).contiguous().clone()
        self.raw_o = raw_opacity.detach().to(**f32).reshape(G).contiguous().clone()
        self.params = [self.means, self.raw_s, self.raw_q, self.raw_o]
        self.exp_avg = [torch.zeros_like(p) for p in self.params]
        self.exp_avg_sq = [torch.zeros_like(p) for p in self.params]
        self.step_count = torch.zeros(1, dtype=torch.int32, device=dev)
        self.lr, self.betas, self.eps = float(lr), betas, float(eps)
        self.w_l1, self.w_ssim = float(w_l1), float(w_ssim)
        settings = cfg.settings()
        settings.raw_activations = 1
        self.batch = GPBatch(frames, sweeps or [SweepSpec(0, G)], G, settings, dev, CapacityHint())
        lay = self.batch.layout
        self.F = lay.n_frames
        P = lay.n_pixels
        self.chunks = [(0, self.F, 0, P, self.batch)]  # one chunk: the whole batch
        self.target = target.detach().to(**f32).reshape(P).contiguous()
        self.image = torch.empty(P, **f32)
        self.d_image = torch.empty(P, **f32)
        self.loss = torch.empty(self.F, **f32)
        self.loss_tmp = torch.empty(int(_lib.lib().us3d_image_loss_tmp_bytes(self.F)) // 8, dtype=torch.float64,
                                    device=dev)
        self.grads = [torch.empty_like(p) for p in self.params]
        self.headroom = headroom
        self.use_graph = use_graph
        self.check_every = max(1, int(check_every))
        self.n_steps = 0
        self.graph = None
        with torch.no_grad():
            st = self.batch.map(self.means, self.raw_s, self.raw_q, self.raw_o).stats
        self._allocate((int(st.n_instances * headroom) + 4096, int(st.n_visible * headroom) + 1024))

    def _allocate(self, caps: Tuple[int, int]) -> None:
        self.k_capacity = min(int(caps[0]), I32_LIMIT)
        self.v_capacity = max(1, min(int(caps[1]), self.batch.layout.n_pairs, I32_LIMIT))
        self.graph = None
        self.workspace = None  # the old workspace goes before the new one is allocated
        self.workspace = new_workspace(self.batch.workspace_bytes(self.k_capacity, self.v_capacity), self.dev)
        self.plan = self.batch.tables.plan(self.workspace, self.k_capacity, self.v_capacity)
        self.fuse_tmp = torch.empty(int(_lib.lib().us3d_gp_loss_tmp_bytes(ctypes.byref(self.plan))) // 8,
                                    dtype=torch.float64, device=self.dev)
        off = self.batch.layout_offsets(self.k_capacity, self.v_capacity)
        self.overflow_ptr = self.workspace.data_ptr() + off.stats + _lib.Stats.overflow.offset
        self.graph = None

    def _body(self) -> None:
        L = _lib.lib()
        s = _lib.stream_ptr(self.dev)
        st = self.batch.settings
        lay = self.batch.layout
        _lib.check(L.us3d_gp_map(self.plan, _lib.ptr(self.means), _lib.ptr(self.raw_s), _lib.ptr(self.raw_q),
                                 _lib.ptr(self.raw_o), None, st, s), "us3d_gp_map")
        g = [_lib.ptr(x) for x in self.grads]
        if self.fused_loss:
            _lib.check(L.us3d_gp_forward_loss(self.plan, st, _lib.ptr(self.image), _lib.ptr(self.target), self.w_l1,
                                              self.w_ssim, 1.0 / lay.n_frames, _lib.ptr(self.loss),
                                              _lib.ptr(self.fuse_tmp), s), "us3d_gp_forward_loss")
            _lib.check(L.us3d_gp_backward_loss(self.plan, st, _lib.ptr(self.image), _lib.ptr(self.target),
                                               _lib.ptr(self.fuse_tmp), g[0], g[1], g[2], g[3], None, s),
                       "us3d_gp_backward_loss")
        else:
            _lib.check(L.us3d_gp_forward(self.plan, st, _lib.ptr(self.image), s), "us3d_gp_forward")
            _lib.check(L.us3d_image_loss(_lib.ptr(self.image), _lib.ptr(self.target),
                                         _lib.ptr(self.batch.tables.pixel_base), lay.n_frames, self.w_l1, self.w_ssim,
                                         1.0 / lay.n_frames, _lib.ptr(self.loss), _lib.ptr(self.d_image),
                                         _lib.ptr(self.loss_tmp), s), "us3d_image_loss")
            _lib.check(L.us3d_gp_backward(self.plan, st, _lib.ptr(self.d_image), g[0], g[1], g[2], g[3], None, s),
                       "us3d_gp_backward")
        b1, b2 = self.betas
        for p, g, m, v in zip(self.params, self.grads, self.exp_avg, self.exp_avg_sq):
            _lib.check(L.us3d_adam_step(_lib.ptr(p), _lib.ptr(g), _lib.ptr(m), _lib.ptr(v), p.numel(),
                                        _lib.ptr(self.step_count), self.overflow_ptr, self.lr, b1, b2, self.eps, s),
                       "us3d_adam_step")
        _lib.check(L.us3d_adam_advance(_lib.ptr(self.step_count), self.overflow_ptr, s), "us3d_adam_advance")

    _capture = SliceSplatTrainer._capture
    overflowed = SliceSplatTrainer.overflowed
    clear_overflow = SliceSplatTrainer.clear_overflow
    step = SliceSplatTrainer.step
    state = SliceSplatTrainer.state
    losses = SliceSplatTrainer.losses
