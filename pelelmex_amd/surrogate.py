"""Learned dt_react map in front of the BDF: dataset, training and loading.

The network maps a cell's state (log10(Y_k + 1e-20), T / 2000, log10(dt)) to its chemical
increments (dY_k, dT) over dt. Training pairs come from the library's GPU BDF (pr_react_box) on
states drawn like the flame-box workload, without external sources (the map learns chemistry;
sources are added exactly in the trust kernel). Forward, backward and Adam run on the library's
own kernels (pr_mlp_*: BF16 matrix-core products with FP32 accumulation); torch only provides
device memory. The trained network is stored as a .npz of float32 arrays and handed to a
Reactor with pr_surrogate_load, after which pr_react_box accepts the cells that pass the FP64
trust check and integrates the others with the BDF.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import time

import numpy as np

from . import workloads
from .reactor import Reactor, ReactorError, lib

HIDDEN = (256, 256, 256, 256)


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _chk(rc, what):
    if rc != 0:
        raise ReactorError(f"{what} failed ({rc}): {lib().pr_last_error().decode()}")


def gemm(M, N, K, A, sam, sak, B, sbk, sbn, bias=None, gelu=False, Cf=None, ldc=0, Cb=None, ldcb=0, Z=None, ldz=0,
         accumulate=False):
    """pr_mlp_gemm on device tensors (A, B bf16)."""
    _chk(lib().pr_mlp_gemm(M, N, K, _ptr(A), sam, sak, _ptr(B), sbk, sbn, _ptr(bias), int(gelu), _ptr(Cf), ldc,
                           _ptr(Cb), ldcb, _ptr(Z), ldz, int(accumulate), _stream()), "pr_mlp_gemm")


class MLP:
    """Dense GELU network with fp32 master parameters and bf16 copies on the device."""

    def __init__(self, dims, device="cuda", seed=0):
        import torch
        self.dims = list(dims)
        self.device = device
        g = torch.Generator().manual_seed(seed)
        self.W, self.b = [], []
        for i, o in zip(self.dims[:-1], self.dims[1:]):
            std = math.sqrt(2.0 / i) if len(self.W) < len(self.dims) - 2 else math.sqrt(1.0 / i)
            self.W.append((torch.randn(i, o, generator=g, dtype=torch.float32) * std).to(device))
            self.b.append(torch.zeros(o, dtype=torch.float32, device=device))
        self._init_state()

    def _init_state(self):
        import torch
        self.Wb = [w.to(torch.bfloat16) for w in self.W]
        self.gW = [torch.zeros_like(w) for w in self.W]
        self.gb = [torch.zeros_like(b) for b in self.b]
        self.m = [torch.zeros_like(p) for p in self.W + self.b]
        self.v = [torch.zeros_like(p) for p in self.W + self.b]
        self.step = 0

    @property
    def nlayers(self):
        return len(self.W)

    def refresh_bf16(self):
        for w, wb in zip(self.W, self.Wb):
            _chk(lib().pr_mlp_to_bf16(w.numel(), _ptr(w), _ptr(wb), _stream()), "pr_mlp_to_bf16")

    def forward(self, X, keep=False):
        """X: bf16 [B, dims[0]] -> fp32 [B, dims[-1]]; keep=True stores what backward needs."""
        import torch
        B = X.shape[0]
        acts, pre = [X], []
        a = X
        for l in range(self.nlayers):
            i, o = self.dims[l], self.dims[l + 1]
            last = l == self.nlayers - 1
            if last:
                out = torch.empty((B, o), dtype=torch.float32, device=X.device)
                gemm(B, o, i, a, i, 1, self.Wb[l], o, 1, bias=self.b[l], Cf=out, ldc=o)
                a = out
            else:
                z = torch.empty((B, o), dtype=torch.float32, device=X.device) if keep else None
                nxt = torch.empty((B, o), dtype=torch.bfloat16, device=X.device)
                gemm(B, o, i, a, i, 1, self.Wb[l], o, 1, bias=self.b[l], gelu=True, Cb=nxt, ldcb=o, Z=z, ldz=o)
                pre.append(z)
                acts.append(nxt)
                a = nxt
        if keep:
            self._acts, self._pre = acts, pre
        return a

    def backward(self, dY):
        """dY: fp32 [B, dims[-1]] gradient of the loss wrt the output; fills gW, gb."""
        import torch
        B = dY.shape[0]
        dZ = dY.contiguous()
        dZb = torch.empty(dZ.shape, dtype=torch.bfloat16, device=dY.device)
        _chk(lib().pr_mlp_to_bf16(dZ.numel(), _ptr(dZ), _ptr(dZb), _stream()), "pr_mlp_to_bf16")
        for l in range(self.nlayers - 1, -1, -1):
            i, o = self.dims[l], self.dims[l + 1]
            a = self._acts[l]
            # gW = A^T dZ  (M = i, N = o, K = B)
            gemm(i, o, B, a, 1, i, dZb, o, 1, Cf=self.gW[l], ldc=o)
            _chk(lib().pr_mlp_column_sum(B, o, _ptr(dZ), o, _ptr(self.gb[l]), 0, _stream()), "pr_mlp_column_sum")
            if l == 0:
                break
            # dA = dZ W^T  (M = B, N = i, K = o), then dZ_prev = dA * gelu'(Z_prev)
            dA = torch.empty((B, i), dtype=torch.float32, device=dY.device)
            gemm(B, i, o, dZb, o, 1, self.Wb[l], 1, o, Cf=dA, ldc=i)
            dZ = torch.empty((B, i), dtype=torch.float32, device=dY.device)
            dZb = torch.empty((B, i), dtype=torch.bfloat16, device=dY.device)
            z = self._pre[l - 1]
            _chk(lib().pr_mlp_gelu_backward(B, i, _ptr(dA), i, _ptr(z), i, _ptr(dZb), i, _ptr(dZ), i, _stream()),
                 "pr_mlp_gelu_backward")

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        params = self.W + self.b
        grads = self.gW + self.gb
        bf = self.Wb + [None] * len(self.b)
        for p, g, m, v, pb in zip(params, grads, self.m, self.v, bf):
            _chk(lib().pr_mlp_adam(p.numel(), _ptr(p), _ptr(g), _ptr(m), _ptr(v), _ptr(pb), lr, b1, b2, eps,
                                   self.step, _stream()), "pr_mlp_adam")

    def mse(self, Y, T):
        """Loss and its gradient (fp32) for outputs Y against targets T."""
        import torch
        dY = torch.empty_like(Y)
        loss = torch.zeros(1, dtype=torch.float32, device=Y.device)
        _chk(lib().pr_mlp_mse(Y.shape[0], Y.shape[1], _ptr(Y), Y.shape[1], _ptr(T), T.shape[1], _ptr(dY), _ptr(loss),
                              _stream()), "pr_mlp_mse")
        return loss, dY

    def flat_params(self):
        out = []
        for w, b in zip(self.W, self.b):
            out += [w.detach().float().cpu().numpy().ravel(), b.detach().float().cpu().numpy().ravel()]
        return np.concatenate(out).astype(np.float32)

    def set_flat_params(self, flat):
        import torch
        off = 0
        for l in range(self.nlayers):
            i, o = self.dims[l], self.dims[l + 1]
            self.W[l].copy_(torch.as_tensor(flat[off:off + i * o].reshape(i, o)))
            off += i * o
            self.b[l].copy_(torch.as_tensor(flat[off:off + o]))
            off += o
        self.refresh_bf16()


def workload_box(mech, seed, dt, nx=64, src_frac=0.0):
    """A training or calibration box of the mechanism's workload: the drm19 flame-like box
    (optionally with sources), or an n-dodecane / syn53 60 atm box (no sources, no mask)."""
    if mech == "dodecane_lu":
        return workloads.dodecane_box(nx=nx, ny=nx, nz=nx, seed=seed, dt=dt, masked=0.0)
    if mech == "drm19":
        return workloads.ch4_air_flame_box(nx=nx, ny=nx, nz=nx, seed=seed, dt=dt, src_frac=src_frac)
    if mech == "syn53":
        return workloads.syn53_chunk(0, nx=nx, ny=nx, nz=nx, seed=seed, dt=dt, masked=0.0)
    raise ValueError(f"no training workload for mechanism {mech}")


def generate_pairs(reactor, seeds, dts, nx=64, device="cuda", log=print, mech="drm19"):
    """(state, result) pairs from the GPU BDF on the mechanism's workload states without external
    sources. Returns device tensors rY0 [N, P], T0 [P], dt [P], rY1 [N, P], T1 [P] and the BDF's
    right-hand side evaluations per cell nfe [P] (its cost)."""
    import torch
    from .reactor import box_tensors
    out = {k: [] for k in ("rY0", "T0", "rH0", "dt", "rY1", "T1", "nfe")}
    for seed, dt in zip(seeds, dts):
        box = workload_box(mech, seed, dt, nx=nx)
        t = box_tensors(box, device)
        rY0, T0 = t["rY"].reshape(reactor.nsp, -1).clone(), t["T"].reshape(-1).clone()
        rH0 = t["rH"].reshape(-1).clone()
        t0 = time.perf_counter()
        status = torch.zeros(T0.numel(), dtype=torch.int32, device=device)
        nfail, _, _ = reactor.react(t["rY"], t["rYsrc"], t["T"], t["rH"], t["rHsrc"], t["FC"], None, dt, 0.0,
                                    status=status)
        torch.cuda.synchronize()
        log(f"  pairs: seed {seed} dt {dt:.2e} cells {T0.numel()} BDF {time.perf_counter() - t0:.2f} s, failed {nfail}")
        ok = status == 0   # a cell whose integration failed gives no pair
        out["rY0"].append(rY0[:, ok])
        out["T0"].append(T0[ok])
        out["rH0"].append(rH0[ok])
        out["dt"].append(torch.full_like(T0[ok], dt))
        out["rY1"].append(t["rY"].reshape(reactor.nsp, -1)[:, ok].clone())
        out["T1"].append(t["T"].reshape(-1)[ok].clone())
        out["nfe"].append(t["FC"].reshape(-1)[ok].clone())
    return {"rY0": torch.cat(out["rY0"], 1), "T0": torch.cat(out["T0"]), "rH0": torch.cat(out["rH0"]),
            "dt": torch.cat(out["dt"]),
            "rY1": torch.cat(out["rY1"], 1), "T1": torch.cat(out["T1"]), "nfe": torch.cat(out["nfe"])}


def targets(pairs):
    """Chemical increments [dY_k (N), dT] per pair, fp64 [P, N + 1]."""
    import torch
    rho0 = pairs["rY0"].sum(0)
    rho1 = pairs["rY1"].sum(0)
    dY = pairs["rY1"] / rho1 - pairs["rY0"] / rho0
    dT = pairs["T1"] - pairs["T0"]
    return torch.cat([dY, dT[None, :]], 0).T.contiguous()


def rate_scales(reactor, pairs, eps_Y=1e-9, eps_T=1e-3):
    """Per-pair output scales dt |f_i(y0)| (/ rho0 for species) from the source-free reactor
    right-hand side (the rates the library's prep kernel evaluates at inference), [P, N + 1],
    and the floors eps [N + 1]."""
    import torch
    N = reactor.nsp
    y0 = torch.cat([pairs["rY0"], pairs["T0"][None, :]], 0).contiguous()
    f = reactor.rhs_batch(y0, torch.zeros_like(pairs["rY0"]), pairs["rH0"].contiguous(),
                          torch.zeros_like(pairs["T0"]), 0.0)
    s = f.abs() * pairs["dt"][None, :]
    s[:N] = s[:N] / pairs["rY0"].sum(0)[None, :]
    eps = torch.full((N + 1,), eps_Y, dtype=torch.float64, device=s.device)
    eps[N] = eps_T
    return s.T.contiguous(), eps


def scaled_targets(reactor, pairs, umax=None):
    """Targets u = d / (scale + eps) and eps; with umax, also the mask of pairs whose every
    |u_i| <= umax (pairs far outside the explicit estimate, such as ignition, are left out)."""
    import torch
    s, eps = rate_scales(reactor, pairs)
    u = targets(pairs) / (s + eps[None, :])
    keep = (u.abs() <= umax).all(1) if umax is not None else torch.ones(u.shape[0], dtype=torch.bool, device=u.device)
    return u, eps, keep


def select_pairs(pairs, keep):
    """The pairs where the boolean device tensor `keep` holds."""
    return {k: (v[:, keep] if v.dim() == 2 else v[keep]) for k, v in pairs.items()}


def train(reactor, pairs, hidden=HIDDEN, epochs=20, batch=4096, lr=1e-3, seed=0, log=print, time_budget=None,
          scaled=True):
    """Trains an MLP on the pairs with the library's kernels; returns (mlp, xnorm, ynorm, history).
    Targets are the increments relative to the rate scales (scaled=True) or the increments
    themselves, z-normalised."""
    import torch
    dev = pairs["T0"].device
    N = reactor.nsp
    raw = reactor.surrogate_features(pairs["rY0"].contiguous(), pairs["T0"], pairs["dt"])
    xmu = raw.double().mean(0)
    xsd = raw.double().std(0).clamp_min(1e-3)
    if scaled:
        tgt, eps, _ = scaled_targets(reactor, pairs)
    else:
        tgt = targets(pairs)
        eps = torch.zeros(tgt.shape[1], dtype=torch.float64, device=dev)
    ymu = tgt.mean(0)
    ysd = tgt.std(0).clamp_min(1e-12)
    xnorm = torch.stack([xmu, xsd]).float().cpu().numpy()
    ynorm = torch.stack([ymu, ysd, eps]).float().cpu().numpy()
    X = ((raw - xmu.float()) / xsd.float()).to(torch.bfloat16).contiguous()
    Y = ((tgt - ymu) / ysd).float().contiguous()
    P = X.shape[0]
    dims = [N + 2, *hidden, N + 1]
    mlp = MLP(dims, device=dev, seed=seed)
    g = torch.Generator(device=dev).manual_seed(seed)
    hist = []
    t0 = time.perf_counter()
    steps_per_epoch = P // batch
    for ep in range(epochs):
        perm = torch.randperm(P, device=dev, generator=g)
        tot = torch.zeros(1, dtype=torch.float32, device=dev)
        for s in range(steps_per_epoch):
            idx = perm[s * batch:(s + 1) * batch]
            xb, yb = X[idx].contiguous(), Y[idx].contiguous()
            out = mlp.forward(xb, keep=True)
            loss, dY = mlp.mse(out, yb)
            mlp.backward(dY)
            lr_t = lr * 0.5 * (1.0 + math.cos(math.pi * (ep * steps_per_epoch + s) / (epochs * steps_per_epoch)))
            mlp.adam(lr=lr_t)
            tot += loss
        torch.cuda.synchronize()
        hist.append(float(tot.item()) / max(steps_per_epoch, 1))
        log(f"  epoch {ep + 1}/{epochs}: mean loss {hist[-1]:.4e} ({time.perf_counter() - t0:.1f} s)")
        if time_budget and time.perf_counter() - t0 > time_budget:
            break
    return mlp, xnorm, ynorm, hist


def save(path, mlp, xnorm, ynorm, **meta):
    np.savez(path, dims=np.asarray(mlp.dims, dtype=np.int32), params=mlp.flat_params(),
             xnorm=np.asarray(xnorm, np.float32), ynorm=np.asarray(ynorm, np.float32),
             **{k: np.asarray(v) for k, v in meta.items()})


def load(path):
    """Dict with dims, params, xnorm, ynorm and the stored metadata (numpy arrays)."""
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def attach(reactor, model):
    """Loads a trained network (dict from load(), or an (mlp, xnorm, ynorm) triple) into a Reactor."""
    if isinstance(model, dict):
        reactor.load_surrogate(model["dims"], model["params"], model["xnorm"], model["ynorm"])
    else:
        mlp, xnorm, ynorm = model
        reactor.load_surrogate(mlp.dims, mlp.flat_params(), xnorm, ynorm)


WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "surrogate_drm19.npz")


def calibrate(reactor_keys, model, seeds=(40100, 40101), target=1e-6, nx=64, device="cuda", log=print,
              grid_rtol=(1e-1, 3e-2, 1e-2, 3e-3, 1e-3, 3e-4, 1e-4), grid_atol=(1e-9, 3e-10, 1e-10, 3e-11, 1e-11),
              mech="drm19", dts=(5e-6,)):
    """Trust tolerances for a trained network. On held-out flame boxes with sources (seeds),
    pr_react_box runs the learned path with each (trust_rtol, trust_atol) of the grid and its
    accepted cells (FC = 1) are compared with a BDF-only call. Among the pairs whose accepted
    cells all stay within `target` relative temperature error on every box (the default is 100x
    below the 1e-4 accuracy bound of the metric, a margin for boxes not seen here), the one that
    takes the most BDF work off is chosen, preferring the smallest trust_atol, then trust_rtol,
    among pairs within 2 % of that work. Returns (best or None, sweep)."""
    import torch
    from .reactor import box_tensors

    def run(t, pr, box, keys, with_model):
        rr = Reactor.create("ReactorCvode", mech, keys)
        rr.init(2, 1)
        if with_model:
            attach(rr, model)
        for k in pr:
            t[k].copy_(pr[k])
        t["FC"].zero_()
        rr.react(t["rY"], t["rYsrc"], t["T"], t["rH"], t["rHsrc"], t["FC"], t["mask"], box["dt"], 0.0)
        torch.cuda.synchronize()
        rr.close()
        return t["T"].reshape(-1).clone(), t["FC"].reshape(-1).clone()

    results = {}
    for i, seed in enumerate(seeds):
        box = workload_box(mech, seed, dts[i % len(dts)], nx=nx, src_frac=0.05)
        t = box_tensors(box, device)
        pr = {k: t[k].clone() for k in ("rY", "T", "rH")}
        T_ref, FC_ref = run(t, pr, box, reactor_keys, False)
        fc = FC_ref.double()
        for rt in grid_rtol:
            for at in grid_atol:
                keys = dict(reactor_keys, **{"surrogate.trust_rtol": rt, "surrogate.trust_atol": at,
                                             "surrogate.audit_frac": 0.0})
                T, FC = run(t, pr, box, keys, True)
                a = FC == 1.0
                e = float(((T[a] - T_ref[a]).abs() / T_ref[a]).max()) if bool(a.any()) else 0.0
                r = results.setdefault((rt, at), {"trust_rtol": rt, "trust_atol": at, "accepted": [], "max_errT": 0.0,
                                                  "bdf_work_share": []})
                r["accepted"].append(float(a.double().mean()))
                r["bdf_work_share"].append(float(fc[a].sum() / fc.sum()))
                r["max_errT"] = max(r["max_errT"], e)
                log(f"  calibration box {seed}: trust_rtol {rt:g} trust_atol {at:g} accepted {float(a.double().mean()):.4f} "
                    f"max dT/T {e:.2e}")
    sweep = []
    for r in results.values():
        r["accepted"] = float(np.mean(r["accepted"]))
        r["bdf_work_share"] = float(np.mean(r["bdf_work_share"]))
        sweep.append(r)
    ok = [r for r in sweep if r["max_errT"] <= target and r["accepted"] > 0]
    best = None
    if ok:
        top = max(r["bdf_work_share"] for r in ok)
        near = [r for r in ok if r["bdf_work_share"] >= 0.98 * top]
        best = min(near, key=lambda r: (r["trust_atol"], r["trust_rtol"]))
    log(f"  calibration: best {best}")
    return best, sweep


def main():
    """Trains a network on the GPU and calibrates its trust tolerances:
    python -m pelelmex_amd.surrogate --out <path.npz> [--mech drm19 --boxes 8 --epochs 20
    --hidden 256,256,256,256 --dts 5e-6]."""
    import argparse
    import json
    import torch
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--boxes", type=int, default=8)
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--nfe-max", type=float, default=200.0)
    ap.add_argument("--umax", type=float, default=20.0)
    ap.add_argument("--seed", type=int, default=40040)
    ap.add_argument("--target", type=float, default=1e-6)
    ap.add_argument("--model", default=None, help="calibrate this trained network instead of training one")
    ap.add_argument("--mech", default="drm19", choices=["drm19", "dodecane_lu", "syn53"])
    ap.add_argument("--hidden", default="256,256,256,256", help="hidden layer widths")
    ap.add_argument("--dts", default="5e-6", help="dt_react values, cycled over the training and calibration boxes")
    args = ap.parse_args()
    log = lambda *a: print(*a, flush=True)
    hidden = tuple(int(x) for x in args.hidden.split(","))
    dts = [float(x) for x in args.dts.split(",")]
    keys = {"ode.rtol": 1e-10, "ode.atol": 1e-10, "cvode.max_order": 4, "cvode.max_substeps": 10000}
    r = Reactor.create("ReactorCvode", args.mech, keys)
    r.init(2, 1)
    dev = torch.device("cuda:0")
    if args.model:
        m = load(args.model)
        best, sweep = calibrate(keys, m, target=args.target, nx=args.nx, device=dev, log=log, mech=args.mech, dts=dts)
        meta = {k: v for k, v in m.items() if k not in ("dims", "params", "xnorm", "ynorm", "trust_rtol", "trust_atol",
                                                       "calibration")}
        meta["calibration"] = json.dumps({"best": best, "sweep": sweep, "target": args.target})
        if best is not None:
            meta["trust_rtol"] = best["trust_rtol"]
            meta["trust_atol"] = best["trust_atol"]
        np.savez(args.out, dims=m["dims"], params=m["params"], xnorm=m["xnorm"], ynorm=m["ynorm"], **meta)
        log(f"saved {args.out}")
        return
    pairs = generate_pairs(r, [args.seed + i for i in range(args.boxes)], [dts[i % len(dts)] for i in range(args.boxes)],
                           nx=args.nx, device=dev, log=log, mech=args.mech)
    _, _, ukeep = scaled_targets(r, pairs, umax=args.umax)
    keep = (pairs["nfe"] <= args.nfe_max) & ukeep
    log(f"pairs {pairs['T0'].numel()}, training on {int(keep.sum())} (BDF cost <= {args.nfe_max} RHS, |u| <= {args.umax})")
    sub = select_pairs(pairs, keep)
    del pairs
    mlp, xn, yn, hist = train(r, sub, hidden=hidden, epochs=args.epochs, log=log)
    best, sweep = calibrate(keys, (mlp, xn, yn), target=args.target, nx=args.nx, device=dev, log=log, mech=args.mech,
                            dts=dts)
    meta = {"loss": np.asarray(hist, np.float64), "train_pairs": int(keep.sum()), "train_seed": args.seed,
            "boxes": args.boxes, "epochs": args.epochs, "nfe_max": args.nfe_max, "umax": args.umax,
            "mech": args.mech, "dts": np.asarray(dts, np.float64),
            "calibration": json.dumps({"best": best, "sweep": sweep, "target": args.target})}
    if best is not None:
        meta["trust_rtol"] = best["trust_rtol"]
        meta["trust_atol"] = best["trust_atol"]
    save(args.out, mlp, xn, yn, **meta)
    log(f"saved {args.out}")


if __name__ == "__main__":
    main()
