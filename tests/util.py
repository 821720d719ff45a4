"""Helpers shared by the tests: seeded inputs, host <-> device copies that
keep numpy strides, error norms."""
import numpy as np


def rand(rng, shape, dtype, kind="uniform"):
    """Seeded col-major (Fortran-ordered) array; re/im i.i.d. U[-1, 1]."""
    shape = tuple(int(s) for s in shape)
    if kind == "uniform":
        x = rng.uniform(-1, 1, shape)
        if np.dtype(dtype).kind == "c":
            x = x + 1j * rng.uniform(-1, 1, shape)
    elif kind == "normal":
        x = rng.standard_normal(shape)
        if np.dtype(dtype).kind == "c":
            x = x + 1j * rng.standard_normal(shape)
    elif kind == "logmag":
        # |z| log-uniform over [1e-3, 1e3], phase uniform on [0, 2 pi)
        r = 10.0 ** rng.uniform(-3, 3, shape)
        th = rng.uniform(0, 2 * np.pi, shape)
        x = r * np.exp(1j * th)
    else:
        raise ValueError(kind)
    x = np.asarray(x).astype(dtype)
    # np.asfortranarray turns a 0-d array into shape (1,); keep scalars 0-d
    return x if x.ndim == 0 else np.asfortranarray(x)


def to_dev(x, device="cuda"):
    """numpy array -> torch device tensor with the same element strides."""
    import torch
    if x.ndim == 0:
        return torch.from_numpy(np.array(x)).to(device)
    if x.flags.f_contiguous:
        t = torch.from_numpy(np.ascontiguousarray(x.T)).to(device)
        return t.permute(*reversed(range(x.ndim)))
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def to_host(t):
    """torch tensor -> numpy array (values; Fortran order if t is col-major)."""
    import torch
    if t.dim() >= 2 and t.stride(0) == 1 and t.stride(-1) != 1:
        return t.permute(*reversed(range(t.dim()))).contiguous().cpu().numpy().T
    return t.contiguous().cpu().numpy()


def rel_err(got, want):
    d = np.linalg.norm((np.asarray(got, dtype=np.complex128) -
                        np.asarray(want, dtype=np.complex128)).ravel())
    n = np.linalg.norm(np.asarray(want, dtype=np.complex128).ravel())
    return d / (n if n > 0 else 1.0)


def extents(labels_list, rng, lo=2, hi=6):
    ext = {}
    for labels in labels_list:
        for ch in labels:
            if ch not in ext:
                ext[ch] = int(rng.integers(lo, hi + 1))
    return ext
