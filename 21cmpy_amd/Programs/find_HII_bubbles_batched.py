"""Batched excursion-set ionization fields: many boxes per launch.

The reference computes one box per ``conv_bubbles`` call
(Programs/find_HII_bubbles.py:17-193) and drives boxes one by one
(Programs/Drives.py:7-14).  Here a set of boxes with their own efficiency
zeta, barrier (redshift) and sigma_min^2 runs as one HIP graph through
``cm21_find_bubbles_batched``: boxes are grouped by (D, box_len); each group
does one batched forward transform, one launch of the pruned columns and rows
passes for all radii and boxes, and one persistent threshold launch that walks
all radii per line group with the ionization mask in LDS.  The per-cell
arithmetic is that of ``conv_bubbles(fused=True)``.

Tile shapes per group come from the library's candidate table; ``autotune``
times them on the actual boxes and the choice is kept in a JSON file
(Parameter_files/batch_tiles_gfx950.json) keyed by the group's shape.
"""
import json
import math
import os

import numpy as np
import torch
from scipy.special import erfc

from .. import _lib
from ..cosmo_files.planck import barrier as planck_barrier
from ..cosmo_files.sigmas import power_law_spectrum
from .find_HII_bubbles import DELTA_C_LINEAR, barrier_denominators, radius_schedule

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MANIFEST = os.path.join(_PKG, "Parameter_files", "batch64_manifest.json")
TILES = os.path.join(_PKG, "Parameter_files", "batch_tiles_gfx950.json")


def linear_growth(z):
    """D(z) = 0.12 * 9 / (1 + z): linear growth anchored at D(8) = 0.12."""
    return 0.12 * 9.0 / (1.0 + float(z))


def make_manifest(nbox=64, seed=0, dims=(96, 128, 160), cell=1.5625, zeta=(10.0, 60.0), redshifts=(7, 8, 9, 10),
                  sigma_min2=(2.0, 6.0)):
    """Parameters of a batch drawn from a fixed generator (NumPy PCG64(seed)):
    box i has GRF seed i, D uniform over dims, box_len = cell * D, zeta and
    sigma_min^2 uniform in their ranges, z uniform over redshifts."""
    rng = np.random.Generator(np.random.PCG64(int(seed)))
    boxes = []
    for i in range(int(nbox)):
        D = int(rng.choice(np.asarray(dims)))
        boxes.append({
            "seed": i,
            "D": D,
            "box_len": float(cell) * D,
            "zeta": float(rng.uniform(*zeta)),
            "z": float(rng.choice(np.asarray(redshifts))),
            "sigma_min2": float(rng.uniform(*sigma_min2)),
        })
    return {"generator": f"numpy.random.PCG64({int(seed)})", "cell_mpc": float(cell),
            "growth": "D(z) = 0.12*9/(1+z)", "input": "linear z=0 density, deltac = 1.686/D(z)",
            "boxes": boxes}


def load_manifest(path=MANIFEST):
    with open(path) as f:
        return json.load(f)


def _tile_key(D, box_len, nbox, fil, LE):
    return f"D{int(D)}_L{float(box_len):g}_B{int(nbox)}_f{int(fil)}_{'lin' if LE else 'nl'}"


def load_tiles(path=TILES):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def box_parameters(spec, LE=True):
    """Barrier, growth factor and mean collapsed fraction of one box from its
    spec: 'deltac' or 'growth' or 'z' (linear growth D(z)), and 'sigma_min2'."""
    if "deltac" in spec:
        deltac = float(spec["deltac"])
    else:
        growth = float(spec["growth"]) if "growth" in spec else linear_growth(spec["z"])
        deltac = DELTA_C_LINEAR / growth
    smin2 = float(spec["sigma_min2"])
    fc_mean = np.float32(erfc(deltac / (math.sqrt(2.0) * math.sqrt(smin2))))
    return deltac, np.float32(deltac / DELTA_C_LINEAR), fc_mean


class BatchedBubbles:
    """Device buffers, plans, tables and graph state for a fixed set of boxes.

    specs: one dict per box with 'D', 'box_len', 'zeta', 'sigma_min2' and the
    barrier ('deltac', 'growth' or 'z'); optional 'sig0' (callable S(R)),
    default the box's synthetic power-law spectrum.  A spec with 'z' and none
    of 'sigma_min2', 'deltac', 'growth' takes the Planck13 barrier and S(R)
    (cosmo_files.planck), as conv_bubbles does.  Boxes are loaded with
    ``set_delta`` (or ``set_deltas``); ``run`` computes every ionized box;
    ``ionized(i)`` is box i's float32 (D,D,D) result (a view, overwritten by
    the next run).
    """

    def __init__(self, specs, fil=1, LE=True, strict_filter=False, graph=True, tiles="auto", stream=None):
        _lib.require_gpu()
        if fil not in (1, 2):
            raise ValueError("the batched loop supports HII_FILTER 1 and 2")
        self.specs = [dict(s) for s in specs]
        for s in self.specs:
            if "sigma_min2" not in s and "deltac" not in s and "growth" not in s:
                # Planck-cosmology mode: the barrier and S(R) from the box's z
                if "z" not in s:
                    raise ValueError("a spec needs 'sigma_min2' with 'deltac', 'growth' or 'z', or 'z' alone")
                b = planck_barrier(float(s["z"]))
                s["deltac"], s["sigma_min2"] = b["deltac"], b["sigma_min2"]
                s.setdefault("sig0", b["sig0"])
        self.LE = bool(LE)
        self.fil = int(fil)
        self.stream = stream
        self.prm = _lib.BubbleParams(filter_type=self.fil, ref_fallthrough=0 if strict_filter else 1,
                                     linear_input=1 if self.LE else 0, update=0, fused=1, graph=1 if graph else 0,
                                     zeta=1.0, box_len=1.0, sphere_r_scale=1.0)
        dev = torch.device("cuda")
        order = {}
        for i, s in enumerate(self.specs):
            order.setdefault((int(s["D"]), float(s["box_len"])), []).append(i)
        self._slot = {}
        self._groups = []
        saved = load_tiles() if tiles == "auto" else (tiles or {})
        sig_cache = {}
        for (D, L), idx in order.items():
            radii, _ = radius_schedule(L, D)
            if (D, L) not in sig_cache:
                sig_cache[(D, L)] = power_law_spectrum(D, L).sig0_discrete
            denoms = np.empty((len(idx), len(radii)), np.float32)
            boxes = (_lib.BatchBox * len(idx))()
            for j, i in enumerate(idx):
                s = self.specs[i]
                deltac, ginv, fc_mean = box_parameters(s, self.LE)
                sig0 = s.get("sig0", sig_cache[(D, L)])
                denoms[j] = barrier_denominators(radii, s["sigma_min2"], sig0)
                boxes[j] = _lib.BatchBox(deltac=float(np.float32(deltac)), growth_inv=float(ginv),
                                         zeta=float(s["zeta"]), fc_mean=float(fc_mean))
                self._slot[i] = (len(self._groups), j)
            delta = torch.empty((len(idx), D, D, D), dtype=torch.float32, device=dev)
            ion = torch.empty_like(delta)
            plan = _lib.plan(D, len(idx))
            key = _tile_key(D, L, len(idx), self.fil, self.LE)
            self._groups.append({"D": D, "L": L, "idx": idx, "delta": delta, "ion": ion, "plan": plan, "key": key,
                                 "radii": radii, "denoms": denoms, "boxes": boxes})
        self._garr = (_lib.BatchGroup * len(self._groups))()
        for k, g in enumerate(self._groups):
            t = saved.get(g["key"], {})
            self._garr[k] = _lib.BatchGroup(
                D=g["D"], nbox=len(g["idx"]), box_len=g["L"], n_radii=len(g["radii"]),
                radii=g["radii"].ctypes.data_as(_lib.c_float_p), denoms=g["denoms"].ctypes.data_as(_lib.c_float_p),
                boxes=g["boxes"], delta=g["delta"].data_ptr(), ionized=g["ion"].data_ptr(), plan=g["plan"].handle,
                tile_columns=int(t.get("columns", -1)), tile_lines=int(t.get("lines", -1)))
        nbytes = _lib.query("cm21_find_bubbles_batched_workspace_bytes", self._garr, len(self._groups), self.prm)
        if nbytes == 0:
            raise ValueError("cm21_find_bubbles_batched_workspace_bytes rejected the groups")
        self.work_bytes = nbytes
        self.work = torch.empty(nbytes, dtype=torch.uint8, device=dev)

    @property
    def groups(self):
        """(D, box_len, box indices) per group, in launch order."""
        return [(g["D"], g["L"], list(g["idx"])) for g in self._groups]

    def tiles(self):
        """Current tile choice per group: {key: (columns index, lines index)}."""
        return {g["key"]: (self._garr[k].tile_columns, self._garr[k].tile_lines)
                for k, g in enumerate(self._groups)}

    def set_tiles(self, choice):
        """choice: {key: (columns, lines)}; -1 = default."""
        for k, g in enumerate(self._groups):
            if g["key"] in choice:
                c, l = choice[g["key"]]
                self._garr[k].tile_columns = int(c)
                self._garr[k].tile_lines = int(l)

    def set_delta(self, i, delta):
        gi, j = self._slot[i]
        self._groups[gi]["delta"][j].copy_(torch.as_tensor(delta, dtype=torch.float32))

    def set_deltas(self, deltas):
        for i, d in enumerate(deltas):
            self.set_delta(i, d)

    def delta(self, i):
        gi, j = self._slot[i]
        return self._groups[gi]["delta"][j]

    def run(self):
        """Compute every ionized box.  With a stream of its own the run waits
        for the caller's current stream (the set_delta copies) and the
        caller's stream waits for the run (its reads of the results)."""
        caller = torch.cuda.current_stream()
        s = caller if self.stream is None else self.stream
        if s != caller:
            s.wait_stream(caller)
        _lib.call("cm21_find_bubbles_batched", self._garr, len(self._groups), self.prm, _lib.ptr(self.work),
                  self.work_bytes, _lib.stream_handle(s))
        if s != caller:
            caller.wait_stream(s)
            for t in self._buffers():
                t.record_stream(s)

    def _buffers(self):
        for g in self._groups:
            yield g["delta"]
            yield g["ion"]
        yield self.work

    def ionized(self, i):
        gi, j = self._slot[i]
        return self._groups[gi]["ion"][j]

    def x_HI(self):
        """Neutral fraction per box (FP64 sums on the device), in spec order."""
        out = np.empty(len(self.specs))
        for g in self._groups:
            D = g["D"]
            s = torch.sum(g["ion"].view(len(g["idx"]), -1), dim=1, dtype=torch.float64).cpu().numpy()
            for j, i in enumerate(g["idx"]):
                out[i] = 1.0 - s[j] / D ** 3
        return out

    def autotune(self, reps=3, save_path=None):
        """Time every tile candidate on the loaded boxes and keep the fastest;
        returns {key: {"columns", "lines", "chosen", "names", "ms"}} (names and
        mean times of the columns candidates, then the lines candidates) and,
        with save_path, merges it into that JSON file."""
        ncand = sum(_lib.query("cm21_batch_tile_count", g["D"], 0) + _lib.query("cm21_batch_tile_count", g["D"], 1)
                    for g in self._groups)
        ms = (_lib.ctypes.c_double * max(ncand, 1))()
        caller = torch.cuda.current_stream()
        stream = self.stream if self.stream is not None else caller
        if stream != caller:
            stream.wait_stream(caller)
        _lib.call("cm21_batch_autotune", self._garr, len(self._groups), self.prm, _lib.ptr(self.work),
                  self.work_bytes, int(reps), ms, _lib.stream_handle(stream))
        res, pos = {}, 0
        lib = _lib.load()
        for k, g in enumerate(self._groups):
            D = g["D"]
            nc = lib.cm21_batch_tile_count(D, 0)
            nl = lib.cm21_batch_tile_count(D, 1)
            names = [lib.cm21_batch_tile_name(D, 0, i).decode() for i in range(nc)] + \
                    [lib.cm21_batch_tile_name(D, 1, i).decode() for i in range(nl)]
            c, l = int(self._garr[k].tile_columns), int(self._garr[k].tile_lines)
            res[g["key"]] = {"columns": c, "lines": l, "chosen": [names[c], names[nc + l]],
                             "names": names, "ms": [float(ms[pos + i]) for i in range(nc + nl)]}
            pos += nc + nl
        if save_path:
            cur = load_tiles(save_path)
            cur.update(res)
            with open(save_path, "w") as f:
                json.dump(cur, f, indent=1, sort_keys=True)
        return res


def conv_bubbles_batched(deltas, specs, fil=1, LE=True, strict_filter=False, graph=False, stream=None):
    """One-shot batched form: ionized boxes (CUDA tensors, copies) for a list
    of density boxes and their specs (see BatchedBubbles)."""
    bb = BatchedBubbles(specs, fil=fil, LE=LE, strict_filter=strict_filter, graph=graph, stream=stream)
    bb.set_deltas(deltas)
    bb.run()
    return [bb.ionized(i).clone() for i in range(len(specs))]
