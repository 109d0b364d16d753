"""On-disk formats on either side of the palette rasterizer.

* Vanilla scene ``.npy`` (the finetune input): a numpy structured array with
  fields ``pw (3)``, ``rot (4)``, ``scale (3)``, ``alpha`` and ``sh (sh_dim)``
  (gsplat/gau_io.py:10-15 ``gsdata_type``; read by gsplat/gau_io.py:130-135 and
  finetune.py:235-246).  Read without unpickling.
* Palette checkpoint ``.npy`` (the finetune output): a
  dict stored by ``np.save`` as a pickled 0-d object array (finetune.py:440-468).  Reading one runs
  the unpickler, so :func:`load_palette_npy` asks the caller to vouch for the
  file (``trusted=True``).  :func:`save_palette_npz` / :func:`load_palette_npz`
  keep the same keys in a pickle-free ``.npz``.
* ``.pply`` + ``.gswp`` (convert_npy_ply/palette_npy_to_pply.py:13-39 for the
  layout, 176-270 for the writers): a geometry-only little-endian binary PLY
  with log scales and logit opacities, and a sidecar with a 32-byte header
  (``GSWP``, version 1, N, splat_dim, num_sh, low_rank, P, Q), the palette, the
  DC-band weights and either the rest bands or rank-1 factors A (N,P), B (N,Q).
  The reference only writes them; :func:`read_pply` and :func:`read_gswp` read
  them back, and :func:`load_scene_files` turns a pair into the arrays the
  rasterizer and the serving loop take.

Palette dicts use the reference's keys: ``pw, rot, scale, alpha, palette``
plus ``shs_w`` (full rank, (N, splat_dim*num_sh) band-major) or
``low_shs_w, high_shs_A, high_shs_B, reshape_P, reshape_Q`` with
``low_rank=True``.  ``splat_dim`` counts all palette rows (black, white and
the chromatic ones).

Command line, as convert_npy_ply/palette_npy_to_pply.py:344-365::

    python -m rgp.formats ckpt.npy [out.pply] [--sidecar x.gswp] [--inspect] [--quiet] --trusted
"""
import argparse
import os
import struct
import sys

import numpy as np

GSWP_MAGIC = b"GSWP"
GSWP_VERSION = 1
_GSWP_HEADER = struct.Struct("<4s7I")  # magic, version, N, splat_dim, num_sh, low_rank, P, Q
_PPLY_FIELDS = ("x", "y", "z", "scale_0", "scale_1", "scale_2", "opacity", "rot_0", "rot_1", "rot_2", "rot_3")
_PPLY_COMMENT = "comment palette-based 3DGS - geometry only; pair with .gswp sidecar"
_LOGIT_EPS = 1e-6   # convert_npy_ply/palette_npy_to_pply.py:77
_SCALE_FLOOR = 1e-20  # convert_npy_ply/palette_npy_to_pply.py:183


def _f32(a):
    return np.ascontiguousarray(np.asarray(a), dtype=np.float32)


# ---------------------------------------------------------------------------
# vanilla scene .npy (structured array)
# ---------------------------------------------------------------------------
def vanilla_dtype(sh_dim):
    """The record type of gsplat/gau_io.py:10-15."""
    return np.dtype([("pw", "<f4", (3,)), ("rot", "<f4", (4,)), ("scale", "<f4", (3,)), ("alpha", "<f4"),
                     ("sh", "<f4", (sh_dim,))])


def save_vanilla_npy(path, pws, rots, scales, alphas, shs):
    """Write a vanilla scene: scales linear, alphas after the sigmoid, shs in the
    layout of gsplat/gau_io.py:94 (the 3 DC values, then the rest band-major)."""
    shs = _f32(shs)
    n = shs.shape[0]
    rec = np.empty(n, dtype=vanilla_dtype(shs.shape[1]))
    rec["pw"], rec["rot"], rec["scale"] = _f32(pws), _f32(rots), _f32(scales)
    rec["alpha"], rec["sh"] = _f32(alphas).reshape(n), shs
    np.save(path, rec, allow_pickle=False)


def load_vanilla_npy(path):
    """Read a vanilla scene into a dict of float32 arrays (pws, rots, scales,
    alphas, shs).  Structured arrays load without the unpickler."""
    rec = np.load(path, allow_pickle=False)
    names = rec.dtype.names or ()
    for k in ("pw", "rot", "scale", "alpha", "sh"):
        if k not in names:
            raise ValueError(f"{path}: not a vanilla scene (field {k!r} missing; fields {names})")
    return {"pws": _f32(rec["pw"]), "rots": _f32(rec["rot"]), "scales": _f32(rec["scale"]),
            "alphas": _f32(rec["alpha"]).reshape(-1), "shs": _f32(rec["sh"])}


# ---------------------------------------------------------------------------
# palette checkpoint
# ---------------------------------------------------------------------------
def save_palette_npy(path, ckpt):
    """The reference's own output format (finetune.py:468): ``np.save`` of the dict."""
    np.save(path, dict(ckpt), allow_pickle=True)


def load_palette_npy(path, trusted=False):
    """Read a palette checkpoint written by ``np.save(fn, dict)``.  That format
    is a pickle, and unpickling a file can run code: pass trusted=True only for
    files from a source you trust (convert_npy_ply/palette_npy_to_pply.py:82-100 loads it the
    same way, without asking)."""
    if not trusted:
        raise PermissionError(f"{path}: a palette .npy is a pickled dict; reading it runs the unpickler. "
                              "Pass trusted=True for a file you trust, or convert it to .npz once.")
    obj = np.load(path, allow_pickle=True)
    if not (isinstance(obj, np.ndarray) and obj.dtype == object and obj.ndim == 0):
        raise ValueError(f"{path}: expected a pickled dict, got an array of {obj.dtype} and shape {obj.shape}")
    ckpt = obj.item()
    if not isinstance(ckpt, dict):
        raise ValueError(f"{path}: holds a {type(ckpt).__name__}, expected dict")
    return ckpt


def checkpoint(pws, rots, scales, alphas, palette, shs=None, dc=None, A=None, B=None):
    """A palette dict with the reference's keys from rasterizer inputs: full rank
    from shs (N, 16K), or low rank from dc (N,K), A (N,P[,1]), B (N,[1,]Q).  The
    format holds rank-1 factors only (P*Q = K*15)."""
    out = {"pw": _f32(pws), "rot": _f32(rots), "scale": _f32(scales), "alpha": _f32(alphas).reshape(-1),
           "palette": _f32(palette)}
    if shs is not None:
        out.update(shs_w=_f32(shs), low_rank=False)
        return out
    A, B = np.asarray(A), np.asarray(B)
    if A.ndim == 3:
        if A.shape[2] != 1 or B.shape[1] != 1:
            raise ValueError(f"factors of rank {A.shape[2]}: the .gswp format holds rank 1; pass the dense shs")
        A, B = A[:, :, 0], B[:, 0, :]
    out.update(low_shs_w=_f32(dc), high_shs_A=_f32(A), high_shs_B=_f32(B), low_rank=True,
               reshape_P=int(A.shape[1]), reshape_Q=int(B.shape[1]))
    return out


def save_palette_npz(path, ckpt):
    """The checkpoint keys in a pickle-free .npz (booleans and ints as 0-d arrays)."""
    np.savez(path, **{k: np.asarray(v) for k, v in ckpt.items()})


def load_palette_npz(path):
    with np.load(path, allow_pickle=False) as z:
        out = {k: z[k] for k in z.files}
    for k in ("low_rank",):
        if k in out:
            out[k] = bool(out[k])
    for k in ("reshape_P", "reshape_Q"):
        if k in out:
            out[k] = int(out[k])
    return out


def palette_info(ckpt):
    """Check a palette dict and return N, splat_dim, num_sh, low_rank (and P, Q
    for low rank).  The checks and error kinds are those of
    convert_npy_ply/palette_npy_to_pply.py:103-173."""
    for k in ("pw", "rot", "scale", "alpha", "palette"):
        if k not in ckpt:
            raise KeyError(f"palette checkpoint is missing {k!r}")
    palette = np.asarray(ckpt["palette"])
    if palette.ndim != 2 or palette.shape[1] != 3:
        raise ValueError(f"palette has shape {palette.shape}, expected (splat_dim, 3)")
    sd = palette.shape[0]
    pw = np.asarray(ckpt["pw"])
    if pw.ndim != 2 or pw.shape[1] != 3:
        raise ValueError(f"pw has shape {pw.shape}, expected (N, 3)")
    n = pw.shape[0]
    for k, shape in (("rot", (n, 4)), ("scale", (n, 3))):
        if np.asarray(ckpt[k]).shape != shape:
            raise ValueError(f"{k} has shape {np.asarray(ckpt[k]).shape}, expected {shape}")
    if np.asarray(ckpt["alpha"]).squeeze().shape != (n,):
        raise ValueError(f"alpha has shape {np.asarray(ckpt['alpha']).shape}, expected ({n},) after squeeze")
    info = {"N": n, "splat_dim": sd, "low_rank": bool(ckpt.get("low_rank", False)), "P": 0, "Q": 0}
    if info["low_rank"]:
        for k in ("low_shs_w", "high_shs_A", "high_shs_B", "reshape_P", "reshape_Q"):
            if k not in ckpt:
                raise KeyError(f"low-rank checkpoint is missing {k!r}")
        P, Q = int(ckpt["reshape_P"]), int(ckpt["reshape_Q"])
        for k, shape in (("low_shs_w", (n, sd)), ("high_shs_A", (n, P)), ("high_shs_B", (n, Q))):
            if np.asarray(ckpt[k]).shape != shape:
                raise ValueError(f"{k} has shape {np.asarray(ckpt[k]).shape}, expected {shape}")
        if (P * Q) % sd:
            raise ValueError(f"P*Q = {P * Q} is not a multiple of splat_dim = {sd}")
        info.update(P=P, Q=Q, num_sh=P * Q // sd + 1)
    else:
        if "shs_w" not in ckpt:
            raise KeyError("full-rank checkpoint is missing 'shs_w'")
        shs = np.asarray(ckpt["shs_w"])
        if shs.ndim != 2 or shs.shape[0] != n or shs.shape[1] % sd:
            raise ValueError(f"shs_w has shape {shs.shape}, expected (N={n}, splat_dim*num_sh) with splat_dim={sd}")
        info["num_sh"] = shs.shape[1] // sd
    return info


# ---------------------------------------------------------------------------
# .pply (geometry-only binary little-endian PLY)
# ---------------------------------------------------------------------------
def _logit(a):
    a = np.clip(a, _LOGIT_EPS, 1.0 - _LOGIT_EPS)
    return np.log(a / (1.0 - a))


def pply_header(n):
    lines = ["ply", "format binary_little_endian 1.0", _PPLY_COMMENT, f"element vertex {n}"]
    lines += [f"property float {f}" for f in _PPLY_FIELDS]
    lines += ["end_header", ""]
    return "\n".join(lines).encode("ascii")


def write_pply(path, ckpt):
    """Geometry of a palette dict: pw as is, log(scale), logit(alpha), rot as is
    (convert_npy_ply/palette_npy_to_pply.py:176-215)."""
    pw, rot, scale = _f32(ckpt["pw"]), _f32(ckpt["rot"]), _f32(ckpt["scale"])
    alpha = _f32(np.asarray(ckpt["alpha"]).squeeze())
    n = pw.shape[0]
    cols = np.empty((n, len(_PPLY_FIELDS)), dtype="<f4")
    cols[:, 0:3] = pw
    cols[:, 3:6] = np.log(np.clip(scale, _SCALE_FLOOR, None)).astype(np.float32)
    cols[:, 6] = _logit(alpha).astype(np.float32)
    cols[:, 7:11] = rot
    with open(path, "wb") as f:
        f.write(pply_header(n))
        f.write(cols.tobytes())


def read_pply(path):
    """Read a little-endian binary PLY whose vertex properties are all floats.
    Returns (dict of property -> float32 column, header comments)."""
    with open(path, "rb") as f:
        data = f.read()
    end = data.find(b"end_header\n")
    if not data.startswith(b"ply\n") or end < 0:
        raise ValueError(f"{path}: not a PLY file")
    lines = data[:end].decode("ascii").split("\n")
    if "format binary_little_endian 1.0" not in lines:
        raise ValueError(f"{path}: only binary little-endian PLY is supported")
    n, props, comments = None, [], []
    for ln in lines:
        tok = ln.split()
        if not tok:
            continue
        if tok[0] == "element":
            if tok[1] != "vertex" or n is not None:
                raise ValueError(f"{path}: expected a single vertex element")
            n = int(tok[2])
        elif tok[0] == "property":
            if tok[1] != "float":
                raise ValueError(f"{path}: property {tok[-1]} is {tok[1]}, only float is supported")
            props.append(tok[2])
        elif tok[0] == "comment":
            comments.append(ln[len("comment "):])
    body = data[end + len(b"end_header\n"):]
    if n is None or len(body) != n * 4 * len(props):
        raise ValueError(f"{path}: payload of {len(body)} bytes does not hold {n} vertices of {len(props)} floats")
    cols = np.frombuffer(body, dtype="<f4").reshape(n, len(props))
    return {p: np.ascontiguousarray(cols[:, i]) for i, p in enumerate(props)}, comments


# ---------------------------------------------------------------------------
# .gswp sidecar
# ---------------------------------------------------------------------------
def write_gswp(path, ckpt, info=None):
    """Palette and weight coefficients.  Full
    rank is split into the DC band (N, splat_dim) and the rest bands
    (N, splat_dim*(num_sh-1)), both band-major (convert_npy_ply/palette_npy_to_pply.py:218-270)."""
    info = info or palette_info(ckpt)
    n, sd, nsh = info["N"], info["splat_dim"], info["num_sh"]
    parts = [_f32(ckpt["palette"])]
    if info["low_rank"]:
        parts += [_f32(ckpt["low_shs_w"]), _f32(ckpt["high_shs_A"]), _f32(ckpt["high_shs_B"])]
    else:
        shs = _f32(ckpt["shs_w"]).reshape(n, nsh, sd)
        parts += [np.ascontiguousarray(shs[:, 0, :]), np.ascontiguousarray(shs[:, 1:, :].reshape(n, -1))]
    with open(path, "wb") as f:
        f.write(_GSWP_HEADER.pack(GSWP_MAGIC, GSWP_VERSION, n, sd, nsh, int(info["low_rank"]), info["P"], info["Q"]))
        for a in parts:
            f.write(a.astype("<f4", copy=False).tobytes())


def read_gswp(path):
    """Read a sidecar into a dict: header fields, palette, low_shs_w and either
    high_shs (N, splat_dim*(num_sh-1)) or high_shs_A (N,P), high_shs_B (N,Q)."""
    with open(path, "rb") as f:
        data = f.read()
    if len(data) < _GSWP_HEADER.size:
        raise ValueError(f"{path}: shorter than the {_GSWP_HEADER.size}-byte header")
    magic, ver, n, sd, nsh, lr, P, Q = _GSWP_HEADER.unpack_from(data)
    if magic != GSWP_MAGIC:
        raise ValueError(f"{path}: bad magic {magic!r}")
    if ver != GSWP_VERSION:
        raise ValueError(f"{path}: version {ver}, only {GSWP_VERSION} is supported")
    if lr not in (0, 1) or sd < 1 or nsh < 1:
        raise ValueError(f"{path}: bad header (low_rank={lr}, splat_dim={sd}, num_sh={nsh})")
    if lr and P * Q != sd * (nsh - 1):
        raise ValueError(f"{path}: P*Q = {P * Q} != splat_dim*(num_sh-1) = {sd * (nsh - 1)}")
    shapes = [("palette", (sd, 3)), ("low_shs_w", (n, sd))]
    shapes += [("high_shs_A", (n, P)), ("high_shs_B", (n, Q))] if lr else [("high_shs", (n, sd * (nsh - 1)))]
    need = _GSWP_HEADER.size + 4 * sum(int(np.prod(s)) for _, s in shapes)
    if len(data) != need:
        raise ValueError(f"{path}: {len(data)} bytes, the header implies {need}")
    out = {"N": n, "splat_dim": sd, "num_sh": nsh, "low_rank": bool(lr), "P": P, "Q": Q}
    off = _GSWP_HEADER.size
    for name, shape in shapes:
        cnt = int(np.prod(shape))
        out[name] = np.frombuffer(data, dtype="<f4", count=cnt, offset=off).reshape(shape).copy()
        off += 4 * cnt
    return out


def sidecar_path_for(pply_path):
    return os.path.splitext(os.fspath(pply_path))[0] + ".gswp"


def export(ckpt, pply_path, sidecar_path=None):
    """Write a palette dict as .pply + .gswp; returns the two paths."""
    pply_path = os.fspath(pply_path)
    sidecar_path = os.fspath(sidecar_path) if sidecar_path else sidecar_path_for(pply_path)
    info = palette_info(ckpt)
    write_pply(pply_path, ckpt)
    write_gswp(sidecar_path, ckpt, info)
    return pply_path, sidecar_path


def convert(npy_path, pply_path=None, sidecar_path=None, verbose=True, trusted=False):
    """convert_npy_ply/palette_npy_to_pply.py:273-322: write <name>/<name>.pply and .gswp next
    to the checkpoint, or at the given paths.  A .npz checkpoint needs no trust."""
    if pply_path is None:
        base = os.path.splitext(os.path.basename(npy_path))[0]
        out_dir = os.path.join(os.path.dirname(os.path.abspath(npy_path)), base)
        os.makedirs(out_dir, exist_ok=True)
        pply_path = os.path.join(out_dir, base + ".pply")
    else:
        os.makedirs(os.path.dirname(os.path.abspath(pply_path)), exist_ok=True)
    sidecar_path = sidecar_path or sidecar_path_for(pply_path)
    ckpt = load_palette_npz(npy_path) if npy_path.endswith(".npz") else load_palette_npy(npy_path, trusted)
    info = palette_info(ckpt)
    export(ckpt, pply_path, sidecar_path)
    if verbose:
        kind = f"rank-1 factors P={info['P']} Q={info['Q']}" if info["low_rank"] else "dense"
        print(f"{npy_path}: N={info['N']:,} splat_dim={info['splat_dim']} num_sh={info['num_sh']} ({kind})")
        print(f"  wrote {pply_path} ({os.path.getsize(pply_path) / 1e6:.2f} MB)")
        print(f"  wrote {sidecar_path} ({os.path.getsize(sidecar_path) / 1e6:.2f} MB)")
    return pply_path, sidecar_path


def load_scene_files(pply_path, sidecar_path=None):
    """A .pply + .gswp pair as rasterizer inputs (float32): pws, rots (as
    stored), scales = exp(log scale), alphas = sigmoid(logit), palette (K,3),
    and shs (N, 16K band-major) or dc (N,K), A (N,P,1), B (N,1,Q).  The scale
    and opacity round trip through log/exp and logit/sigmoid is not exact: a
    few ulp for scales, and opacities are clipped to [1e-6, 1 - 1e-6]."""
    cols, _ = read_pply(pply_path)
    for f in _PPLY_FIELDS:
        if f not in cols:
            raise ValueError(f"{pply_path}: property {f!r} missing")
    sc = read_gswp(sidecar_path or sidecar_path_for(pply_path))
    n = sc["N"]
    if cols["x"].shape[0] != n:
        raise ValueError(f"{pply_path} holds {cols['x'].shape[0]} vertices, the sidecar {n}")
    st = lambda names: np.ascontiguousarray(np.stack([cols[c] for c in names], axis=1), dtype=np.float32)
    logit = cols["opacity"].astype(np.float64)
    out = {"pws": st(("x", "y", "z")), "rots": st(("rot_0", "rot_1", "rot_2", "rot_3")),
           "scales": np.exp(st(("scale_0", "scale_1", "scale_2")).astype(np.float64)).astype(np.float32),
           "alphas": (1.0 / (1.0 + np.exp(-logit))).astype(np.float32), "palette": sc["palette"],
           "K": sc["splat_dim"], "num_sh": sc["num_sh"], "low_rank": sc["low_rank"]}
    if sc["low_rank"]:
        out.update(dc=sc["low_shs_w"], A=sc["high_shs_A"].reshape(n, sc["P"], 1),
                   B=sc["high_shs_B"].reshape(n, 1, sc["Q"]), P=sc["P"], Q=sc["Q"])
    else:
        out["shs"] = np.ascontiguousarray(np.concatenate([sc["low_shs_w"], sc["high_shs"]], axis=1))
    return out


def inspect(path, trusted=False):
    ckpt = load_palette_npz(path) if path.endswith(".npz") else load_palette_npy(path, trusted)
    info = palette_info(ckpt)
    print(f"{path}: keys {sorted(ckpt)}")
    print(f"  N={info['N']:,} splat_dim={info['splat_dim']} num_sh={info['num_sh']} low_rank={info['low_rank']}")
    if info["low_rank"]:
        dense = info["N"] * info["splat_dim"] * (info["num_sh"] - 1)
        print(f"  P={info['P']} Q={info['Q']}: {dense:,} dense rest-band floats as "
              f"{info['N'] * (info['P'] + info['Q']):,} factor floats")
    print(np.asarray(ckpt["palette"]).round(4))
    return info


def main(argv=None):
    ap = argparse.ArgumentParser(description="palette checkpoint -> .pply + .gswp")
    ap.add_argument("input", help="palette checkpoint (.npy dict, or .npz with the same keys)")
    ap.add_argument("output", nargs="?", default=None, help="output .pply path")
    ap.add_argument("--sidecar", default=None, help="sidecar .gswp path (default: next to the .pply)")
    ap.add_argument("--inspect", action="store_true", help="describe the checkpoint; write nothing")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--trusted", action="store_true", help="allow unpickling a .npy checkpoint")
    a = ap.parse_args(argv)
    if a.inspect:
        inspect(a.input, a.trusted)
    else:
        convert(a.input, a.output, a.sidecar, verbose=not a.quiet, trusted=a.trusted)
    return 0


if __name__ == "__main__":
    sys.exit(main())
