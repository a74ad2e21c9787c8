"""Calls of the fused projection and its adjoint straight through the C ABI
(libatdf.so via ctypes), for GPU tests that need f32 gradients of a BF16
grid or the raw entry points."""
import ctypes as C

import torch

from adaptivetensordf_amd import capi, native


class Fused:
    """Interval list, packed tree and field descriptor of one (tree, rays,
    grid, decoder) combination."""

    def __init__(self, tree, o_t, d_t, K, grid, mlp, C_, G=11, H=64, jitter_mode=2, jitter_seed=0, bf16_decoder=False,
                 intervals=None):
        self.tree, self.o, self.d = tree, o_t, d_t
        self.R = o_t.shape[0]
        sp = native().SamplingParams()
        sp.max_samples_per_node = K
        sp.jitter_mode = jitter_mode
        sp.jitter_seed = jitter_seed
        self.sp_native = sp
        self.iv = intervals if intervals is not None else tree.CreateIntervalsForRays(native().RayList(o_t, d_t), sp)
        self.grid, self.mlp = grid, mlp
        bf16_grid = grid.dtype == torch.bfloat16
        self.field = capi.Field(C_, H, G, capi.GRID_BF16 if bf16_grid else capi.GRID_F32, 2.0, 0, grid.data_ptr(),
                                grid.shape[0], tree.node_to_grid.data_ptr(), mlp.data_ptr(),
                                capi.DECODER_BF16 if bf16_decoder else capi.DECODER_F32, 0)
        self.sp = capi.Sampling(K, 256, jitter_mode, 0, jitter_seed, None, None, None)
        iv = self.iv
        self.civ = capi.Intervals(iv.node.data_ptr(), iv.tmin.data_ptr(), iv.tmax.data_ptr(), iv.count.data_ptr(),
                                  iv.first.data_ptr(), iv.ray.data_ptr(), iv.num_intervals)
        self.packed = tree.PackedTree()

    def args(self):
        return (capi.ptr(self.packed), self.tree.NumNodes(), capi.ptr(self.o), capi.ptr(self.d), self.R,
                C.byref(self.sp), capi.ptr(self.iv.ray_offset), C.byref(self.civ), C.byref(self.field))

    def forward(self):
        P = torch.empty(self.R, device=self.o.device)
        capi.call("atdf_project_fwd", *self.args(), capi.ptr(P), capi.stream_ptr())
        return P

    def workspace(self, staging=0):
        """The adjoint's workspace; staging > 0 adds that many bytes for its
        two-pass form (staging_bytes() covers the whole list)."""
        n = capi.lib().atdf_project_bwd_workspace_bytes(self.R, self.iv.num_intervals, C.byref(self.field))
        return torch.empty(n + staging, dtype=torch.uint8, device=self.o.device)

    def staging_bytes(self):
        return capi.lib().atdf_project_bwd_staging_bytes(self.iv.num_intervals, self.iv.num_samples,
                                                         C.byref(self.field))

    def backward(self, dP, staging=0):
        """(grad_grid f32 [slots, G, G, G, C], grad_mlp f32) for dL/dP."""
        gshape = tuple(self.grid.shape)
        gg = torch.zeros(gshape, dtype=torch.float32, device=self.o.device)
        gm = torch.zeros_like(self.mlp)
        ws = self.workspace(staging)
        dP = dP.contiguous().float()
        capi.call("atdf_project_bwd", *self.args(), capi.ptr(dP), capi.ptr(gg), capi.ptr(gm), capi.ptr(ws),
                  ws.numel(), capi.stream_ptr())
        return gg, gm
