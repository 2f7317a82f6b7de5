"""Decision records of the hard-threshold filters (dft2_process_batch_decisions): runs a
clip with them and decodes them into per-coefficient decisions [..., T, 16, 9], the
layout of the reference's spectrum (t, ky, kx) (nvrtc_source/kernels.cu:45).  Parity
instrumentation: a checker replays them to show where another arithmetic's hard
decisions differ.

Record layout (include/dfttest2_mi355x.h): per output frame and block, (radius+1)*4
64-bit masks [g][slot]; bit `lane` is the decision on the coefficient middle-stage lane
`lane` owns at (g, slot).  Coefficients (kt, ky, kx) with kx = 0 or 8 and kt > radius are
not decided by any lane: the C2R convention takes only the real part of those columns,
which their conjugates (T - kt, 16 - ky, kx) carry, so they take their conjugates'
decisions.
"""

import numpy as np

from .dfttest2 import plane_job


def lane_positions(t_len):
    """[(g, slot, lane, kt, ky, kx)] for every recorded bit."""
    r = t_len // 2
    out = []
    for g in range(r + 1):
        for slot in range(4):
            e, k = slot >> 1, slot & 1
            for lane in range(64):
                pp, q = lane >> 3, lane & 7
                ky = (pp if e == 0 else (16 - pp) % 16) if pp else (0 if e == 0 else 8)
                if q:
                    pos = (g, ky, q) if k == 0 else ((t_len - g) % t_len, (16 - ky) % 16, q)
                else:
                    pos = (g, ky, 0 if k == 0 else 8)
                out.append((g, slot, lane) + pos)
    return out


def decode(masks, t_len):
    """masks [..., radius+1, 4] uint64 -> decisions [..., T, 16, 9] bool.  Raises if a
    coefficient decided twice (temporal pair 0 holds each kx = 1..7 coefficient in two
    slots) was decided differently, or one is left undecided."""
    r = t_len // 2
    lead = masks.shape[:-2]
    dec = np.zeros(lead + (t_len, 16, 9), bool)
    seen = np.zeros((t_len, 16, 9), int)
    for g, slot, lane, kt, ky, kx in lane_positions(t_len):
        bit = ((masks[..., g, slot] >> np.uint64(lane)) & np.uint64(1)).astype(bool)
        if seen[kt, ky, kx]:
            if not np.array_equal(dec[..., kt, ky, kx], bit):
                raise AssertionError(f"coefficient {(kt, ky, kx)} decided differently in its two slots")
        dec[..., kt, ky, kx] = bit
        seen[kt, ky, kx] += 1
    for kt in range(r + 1, t_len):
        for ky in range(16):
            for kx in (0, 8):
                dec[..., kt, ky, kx] = dec[..., (t_len - kt) % t_len, (16 - ky) % 16, kx]
                seen[kt, ky, kx] += 1
    if (seen == 0).any():
        raise AssertionError("undecided coefficients in the record layout")
    return dec


def is_generic(p) -> bool:
    """Plans the library runs on its generic block path (block_size != 16 or radius > 3)."""
    return p.block_size != 16 or p.radius > 3


def decode_generic(masks, t_len, block_size):
    """Generic-path record [..., words] uint64 -> decisions [..., T, B, B/2+1]: bit i of the
    mask sequence is coefficient i in [kt][ky][kx] order."""
    n_coef = t_len * block_size * (block_size // 2 + 1)
    bits = (masks[..., :, None] >> np.arange(64, dtype=np.uint64)) & np.uint64(1)
    flat = bits.reshape(masks.shape[:-1] + (-1,))[..., :n_coef].astype(bool)
    return flat.reshape(masks.shape[:-1] + (t_len, block_size, block_size // 2 + 1))


def run_with_decisions(plan, x, p):
    """Filters the clip x [N, H, W] (numpy) on the GPU through the decision-recording entry;
    returns (output, decisions per output frame [ny, nx, T, B, B/2+1])."""
    import torch

    n, h, w = x.shape
    B = p.block_size
    ps = lambda size: size + ((B - size % B) if size % B else 0) + max(B - p.block_step, p.block_step) * 2  # noqa: E731
    ny, nx = (ps(h) - B) // p.block_step + 1, (ps(w) - B) // p.block_step + 1
    t_len = p.t_len
    words = (t_len * B * (B // 2 + 1) + 63) // 64 if is_generic(p) else (t_len // 2 + 1) * 4
    src = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    dst = torch.empty_like(src)
    n_masks = n * ny * nx * words
    rec = torch.zeros(n_masks, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    plan.process_batch_decisions([plane_job(src, dst)], rec.data_ptr(), n_masks, None)
    out = dst.cpu().numpy()
    masks = rec.cpu().numpy().view(np.uint64)
    if is_generic(p):
        return out, list(decode_generic(masks.reshape(n, ny, nx, words), t_len, B))
    return out, list(decode(masks.reshape(n, ny, nx, t_len // 2 + 1, 4), t_len))
