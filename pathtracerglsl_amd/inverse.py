"""Differentiable media inverse render over the adjoint pair (pt_forward / pt_backward).

Parameters theta = [sigma_s RGB, sigma_a RGB, HG g, albedo x n] (the adjoint's layout, see
include/ptmi.h pt_adjoint_params) live in the material records of a compiled scene; each step
writes them with pt_scene_update_materials, renders the primal image, forms the L2 loss
against a target image and its adjoint image dL/dI = 2 (I - I_target) / N, and replays the
paths with pt_backward for dL/dtheta. The primal and the replayed paths use different
resetSalt streams, so the loss gradient is not correlated with the primal noise.
"""
import numpy as np

from . import abi


def write_params(materials, theta, medium_material, albedo_materials):
    """Material records with theta written into the parameter fields."""
    m = np.array(materials, dtype=abi.MATERIAL_DTYPE, copy=True)
    theta = np.asarray(theta, np.float64)
    if medium_material >= 0:
        m[medium_material]["sigma_s"][:3] = theta[0:3]
        m[medium_material]["sigma_a"][:3] = theta[3:6]
        m[medium_material]["hg_g"] = theta[6]
    for k, mat in enumerate(albedo_materials):
        m[mat]["albedo"][:3] = theta[7 + k]
    return m


def clamp_params(theta):
    """Keep theta in the adjoint's domain: sigma_s > 0, sigma_a >= 0, |g| < 1, albedo in (0, 1)."""
    t = np.array(theta, np.float64, copy=True)
    t[0:3] = np.maximum(t[0:3], 1e-3)
    t[3:6] = np.maximum(t[3:6], 0.0)
    t[6] = float(np.clip(t[6], -0.95, 0.95))
    t[7:] = np.clip(t[7:], 1e-3, 0.999)
    return t


class InverseRender:
    """Adjoint-driven fit of theta to a target image on one compiled scene."""

    def __init__(self, engine, compiled, medium_material, albedo_materials, spp,
                 reduce_mode=abi.REDUCE_DETERMINISTIC):
        self.engine = engine
        self.comp = compiled
        self.spp = int(spp)
        self.medium = int(medium_material)
        self.albedos = [int(a) for a in albedo_materials]
        self.adj = abi.PtAdjointParams.make(self.medium, self.albedos, reduce_mode)
        self.n = 7 + len(self.albedos)
        self.base = np.array(compiled.scene.materials, dtype=abi.MATERIAL_DTYPE, copy=True)
        engine.load_renderer(compiled)

    def set_params(self, theta):
        self.engine.update_materials(write_params(self.base, theta, self.medium, self.albedos))

    def render(self, theta, salt):
        """Primal image [H][W][4] at theta, resetSalt salt."""
        self.set_params(theta)
        self.engine.forward(self.spp, self.comp.params(salt))
        return self.engine.read_export()

    def gradient(self, theta, dL_dimage, salt):
        """dL/dtheta for the adjoint image, replaying the paths of resetSalt salt."""
        self.set_params(theta)
        g = self.engine.backward(self.spp, self.adj, dL_dimage, self.comp.params(salt))
        return g[:self.n].astype(np.float64)

    def loss_and_grad(self, theta, target, step=0):
        img = self.render(theta, 2 * step + 1)
        diff = img[..., :3].astype(np.float64) - target[..., :3]
        npx = diff.shape[0] * diff.shape[1]
        loss = float((diff * diff).sum() / npx)
        dL = np.zeros(img.shape, np.float32)
        dL[..., :3] = 2.0 * diff / npx
        return loss, self.gradient(theta, dL, 2 * step + 2)

    def fit(self, theta0, target, steps=200, lr=0.02, betas=(0.9, 0.999), eps=1e-8, log=None):
        """Adam on theta (learning rate relative to the initial magnitude of each parameter)."""
        theta = clamp_params(theta0)
        scale = np.maximum(np.abs(theta), 1e-2)
        m = np.zeros(self.n)
        v = np.zeros(self.n)
        history = []
        for it in range(steps):
            loss, g = self.loss_and_grad(theta, target, it)
            g = g * scale                                    # Adam in units of the initial magnitude
            m = betas[0] * m + (1 - betas[0]) * g
            v = betas[1] * v + (1 - betas[1]) * g * g
            mh = m / (1 - betas[0] ** (it + 1))
            vh = v / (1 - betas[1] ** (it + 1))
            theta = clamp_params(theta - lr * scale * mh / (np.sqrt(vh) + eps))
            history.append(loss)
            if log is not None and (it % 20 == 0 or it == steps - 1):
                log(f"[inverse] step {it} loss {loss:.6g}")
        return theta, history


def relative_error(a, b):
    """||a - b|| / ||b|| (gradient error against a reference)."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def param_error(theta, target):
    """Mean relative parameter error."""
    theta, target = np.asarray(theta, np.float64), np.asarray(target, np.float64)
    return float(np.mean(np.abs(theta - target) / np.maximum(np.abs(target), 1e-6)))



def main(argv=None):
    """The differentiable media inverse render: hidden target (seeded) rendered at res^2 x
    target_spp, then Adam from the perturbed init with the adjoint pair."""
    import argparse
    import json
    import time
    from . import compiler, scenes
    from .engine import Engine
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--res", type=int, default=512)
    ap.add_argument("--spp", type=int, default=64, help="spp of the target image and of each step")
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--lr", type=float, default=0.02)
    ap.add_argument("--depth", type=int, default=32)
    ap.add_argument("--seed", type=int, default=2024)
    a = ap.parse_args(argv)
    target, init = scenes.inverse_media_params(a.seed)
    strategy = compiler.RenderStrategy("mis", a.depth, 3, 0.95, 64)
    comp = compiler.compile_scene(scenes.inverse_media_scene(target), strategy, a.res, a.res)
    med, alb = scenes.inverse_adjoint_materials(comp)
    with Engine(0) as eng:
        inv = InverseRender(eng, comp, med, alb, a.spp)
        tgt = inv.render(target, 0x7a46e7)
        t0 = time.perf_counter()
        theta, hist = inv.fit(init, tgt, steps=a.steps, lr=a.lr, log=lambda m: print(m, flush=True))
        dt = time.perf_counter() - t0
    out = {"res": a.res, "spp": a.spp, "steps": a.steps, "lr": a.lr, "seconds": round(dt, 2),
           "ms_per_step": round(1e3 * dt / max(a.steps, 1), 2),
           "loss_first": hist[0], "loss_last10": float(np.mean(hist[-10:])),
           "param_error_init": param_error(init, target), "param_error_final": param_error(theta, target),
           "target": [round(float(v), 4) for v in target], "recovered": [round(float(v), 4) for v in theta]}
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
