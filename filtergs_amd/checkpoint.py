"""LoG checkpoints as FilterGS inputs (apps/render.py:101-136, 340-370;
LoG/utils/command.py:12-16).

A LoG checkpoint is a state dict, bare (``*_wotrain.pth``) or wrapped as
``{'state_dict': ..., 'global_iterations': ...}`` (LoG/utils/trainer.py:626-640).
The keys used here:

* ``gaussian.xyz`` [N,3], ``gaussian.scaling`` [N,3] (log scales),
  ``gaussian.rotation`` [N,4] (raw quaternions), ``gaussian.opacity`` [N,1]
  (logits), ``gaussian.colors`` [N,3] (SH DC), ``gaussian.shs`` [N,M-1,3];
* ``tree.node_index`` [N] (-1 at a leaf), ``tree.index_parent`` [N] (-1 at a
  root), ``tree.depth`` [N], ``tree.ancestor_path`` [N,L] (optional: it can be
  rebuilt from the parent links, or the cut can use the links directly).

Files are read with ``torch.load(weights_only=True)``: tensors and plain
containers only, nothing from the file is executed.
"""
from __future__ import annotations

import torch

from . import lod
from .filtergs import FilterGS

_GAUSSIAN_KEYS = ("xyz", "scaling", "rotation", "opacity", "colors", "shs")


def load_statedict(ckptname, map_location="cpu"):
    """The state dict of a LoG checkpoint file (LoG/utils/command.py:12-16)."""
    sd = torch.load(ckptname, map_location=map_location, weights_only=True)
    if isinstance(sd, dict) and "state_dict" in sd:
        sd = sd["state_dict"]
    if not isinstance(sd, dict):
        raise ValueError(f"{ckptname}: not a state dict")
    return sd


def tree_tensors(state_dict, device="cuda", rebuild_ancestor=False):
    """The LoD-cut inputs of the tree: dict(node_index, index_parent, depth,
    ancestor_path) as int32 on ``device``.  ancestor_path comes from the
    checkpoint, or is rebuilt from the parent links on the GPU when it is
    missing or ``rebuild_ancestor`` is set (apps/ancestor.py)."""
    for key in ("tree.node_index", "tree.index_parent", "tree.depth"):
        if key not in state_dict:
            raise KeyError(f"Missing '{key}' in checkpoint state_dict")
    out = {k: state_dict[f"tree.{k}"].to(device=device, dtype=torch.int32).contiguous()
           for k in ("node_index", "index_parent", "depth")}
    if rebuild_ancestor or not any(k in state_dict for k in ("tree.ancestor_path", "tree.ancestor")):
        sd = {"tree.index_parent": out["index_parent"], "tree.depth": out["depth"]}
        for k in ("tree.ancestor_path", "tree.ancestor"):
            if k in state_dict:
                sd[k] = state_dict[k]
                break
        ap, _ = lod.rebuild_ancestor_path(sd, device=device)
        out["ancestor_path"] = ap.to(torch.int32).contiguous()
    else:
        key = "tree.ancestor_path" if "tree.ancestor_path" in state_dict else "tree.ancestor"
        out["ancestor_path"] = state_dict[key].to(device=device, dtype=torch.int32).contiguous()
    return out


def filtergs_from_statedict(state_dict, device="cuda", sh_degree=None, active_sh_degree=None,
                            xyz_scale=1.0):
    """A FilterGS holding the checkpoint's internal parameterisation, as
    load_filtergs_from_log builds it (apps/render.py:101-123).  sh_degree
    defaults to the one the SH buffer holds."""
    missing = [k for k in _GAUSSIAN_KEYS[:5] if f"gaussian.{k}" not in state_dict]
    if missing:
        raise KeyError(f"Missing gaussian keys in checkpoint state_dict: {missing}")
    shs = state_dict.get("gaussian.shs")
    rest = int(shs.shape[1]) if shs is not None and shs.dim() == 3 else 0
    if sh_degree is None:
        sh_degree = int(round((rest + 1) ** 0.5)) - 1
    fg = FilterGS(sh_degree=int(sh_degree), xyz_scale=float(xyz_scale)).to(device)
    for k in _GAUSSIAN_KEYS:
        key = f"gaussian.{k}"
        if key in state_dict:
            fg.register_buffer(k, state_dict[key].to(device=device, dtype=torch.float32).contiguous())
    fg.active_sh_degree = int(sh_degree if active_sh_degree is None else active_sh_degree)
    fg.eval()
    return fg
