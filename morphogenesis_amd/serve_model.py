"""Queueing model of the rolling-sweep request loop (csrc/mg_server.cpp).

The loop admits requests FIFO at the start of a sweep step while their keys fit
under the key cap; a step costs (active keys) x (GPU time per key) / (steps per
sweep); a request finishes after `steps` steps. Given the per-key cost that
the scan kernels reach, the model gives the latency percentiles the queue
itself imposes, so a measured p99 can be told apart from loop overhead, and
the cap that suits an arrival process can be chosen without a GPU.

CLI: python -m morphogenesis_amd.serve_model --key-ms 2.1 --cap 96 --cap 192
"""
import argparse
import json

import numpy as np


def arrivals(rate, duration, k_min=1, k_max=32, keys_per_lookup=3, seed=0):
    """Poisson arrivals over `duration` seconds: a list of (time, keys), each
    request carrying keys_per_lookup x U{k_min..k_max} keys."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        out.append((t, keys_per_lookup * int(rng.integers(k_min, k_max + 1))))


POLICIES = ("fifo", "skip", "smallest")


def simulate(reqs, cap, key_s, steps=4, step_overhead_s=0.0, policy="fifo"):
    """Latencies (seconds, in completion order) of `reqs` under an admission
    policy: "fifo" is the loop's (the head of the queue blocks those behind
    it), "skip" admits later requests that fit past a blocked head, "smallest"
    admits the smallest waiting requests first. A request larger than `cap` is
    refused by the server; here it raises ValueError."""
    if any(k > cap for _, k in reqs):
        raise ValueError("a request exceeds the key cap")
    if policy not in POLICIES:
        raise ValueError(f"unknown policy {policy!r}")
    now, i, active, waiting, lat = 0.0, 0, [], [], []
    while i < len(reqs) or active or waiting:
        if not active and not waiting and reqs[i][0] > now:
            now = reqs[i][0]  # idle until the next arrival
        while i < len(reqs) and reqs[i][0] <= now:
            waiting.append(reqs[i])
            i += 1
        n_act = sum(a[1] for a in active)
        if policy == "smallest":
            waiting.sort(key=lambda r: r[1])
        j = 0
        while j < len(waiting):
            if n_act + waiting[j][1] <= cap:
                t, k = waiting.pop(j)
                active.append([steps, k, t])
                n_act += k
            elif policy == "fifo":
                break
            else:
                j += 1
        now += n_act * key_s / steps + step_overhead_s
        for a in active:
            a[0] -= 1
        lat.extend(now - a[2] for a in active if a[0] == 0)
        active = [a for a in active if a[0] > 0]
    return np.array(lat)


def percentiles(lat):
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--key-ms", type=float, action="append",
                    help="GPU time per key over the whole matrix (repeatable)")
    ap.add_argument("--cap", type=int, action="append", help="key cap (repeatable)")
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--rate", type=float, default=5.0)
    ap.add_argument("--duration", type=float, default=2000.0)
    ap.add_argument("--k-max", type=int, default=32)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--policy", choices=POLICIES, default="fifo")
    args = ap.parse_args()
    reqs = arrivals(args.rate, args.duration, k_max=args.k_max, seed=args.seed)
    for key_ms in args.key_ms or [2.1]:
        for cap in args.cap or [96]:
            lat = simulate(reqs, cap, key_ms * 1e-3, args.steps, policy=args.policy)
            print(json.dumps({"key_ms": key_ms, "cap": cap, "steps": args.steps,
                              "policy": args.policy,
                              "requests": len(reqs), **percentiles(lat)}))


if __name__ == "__main__":
    main()
