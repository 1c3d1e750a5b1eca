"""Train the viscosity surrogate on a run of its own and store the weights.

The training run is the benchmark box (geoflac_amd/configs/layers4000x1000.inp) with
a different seed of the layer-interface offsets and cohesion factors than the runs
that use the weights (seed 1; the benchmark uses seed 0), advanced in FP32 arithmetic.
The solver's own viscosity inputs and outputs are sampled at the given steps (one
sample per element each time), the network is trained on the device by the
library's forward and backward kernels (gf_surrogate_train), and the weights with
their normalisation and gate tables go to an .npz file:

    w     float32[4737]  network parameters
    cfg   float32[66]    input normalisation and per-phase gate ranges
    meta  JSON string    the run, sample steps, training settings, losses

usage: python -m geoflac_amd.train_surrogate [--out PATH] [--seed 1] [--steps 0 10 100 300]
"""
import argparse
import json
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "weights", "visc_mlp_layers4000x1000.npz")


def train(config="layers4000x1000", seed=1, sample_steps=(0, 10, 100, 300), train_steps=3000, batch=8192,
          lr=2e-3, out=DEFAULT_OUT, interface_noise=2.0, coh_noise=0.01):
    import numpy as np
    from . import abi, gpu, inp, model
    p = inp.read_inp(os.path.join(HERE, "configs", config + ".inp"))
    gp, st = model.build_box(p, interface_noise=interface_noise, coh_noise=coh_noise, seed=seed)
    E = st.nelem
    cap = len(sample_steps) * E
    t0 = time.time()
    with gpu.Model(gp, st, precision=abi.PREC_MIXED) as m:
        model.device_setup(m)
        done, n = 0, 0
        for s in sorted(sample_steps):
            if s > done:
                m.step(s - done)
                done = s
            n = m.surrogate_sample(cap)
        l0, l1 = m.surrogate_train(train_steps, batch=batch, lr=lr, seed=1)
        w, cfg = m.surrogate_weights()
    meta = {"config": config, "seed": seed, "precision": "fp32 arithmetic", "sample_steps": list(sample_steps),
            "samples": n, "train_steps": train_steps, "batch": batch, "lr": lr, "loss_first": l0, "loss_last": l1,
            "interface_noise_km": interface_noise, "cohesion_noise": coh_noise,
            "seconds": round(time.time() - t0, 1)}
    os.makedirs(os.path.dirname(out), exist_ok=True)
    np.savez(out, w=w, cfg=cfg, meta=np.array(json.dumps(meta)))
    return meta


def load(path=DEFAULT_OUT):
    """(w, cfg, meta) of a stored network."""
    import numpy as np
    z = np.load(path, allow_pickle=False)
    return z["w"], z["cfg"], json.loads(str(z["meta"]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--steps", type=int, nargs="+", default=[0, 10, 100, 300])
    ap.add_argument("--train-steps", type=int, default=3000)
    args = ap.parse_args()
    meta = train(seed=args.seed, sample_steps=tuple(args.steps), train_steps=args.train_steps, out=args.out)
    print(json.dumps(meta), flush=True)


if __name__ == "__main__":
    main()
