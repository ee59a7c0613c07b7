"""Synthetic inputs for the benchmark workloads (host side, numpy).

Each function returns the model parameters (the float vector shared by the
model libraries and the checker) and a seeded initial population.

circles   : spatial-3D circles, N agents uniform in [0, W)^3, W = cbrt(N),
            r = 2.0, repulse 0.05 (examples/cpp/circles_spatial3D/src/main.cu:5-54)
boids     : spatial-3D boids with wrap-around, N agents uniform in a 3.2-wide
            cube, |v| = 0.05, r = 0.05 -> 64^3 bins
            (examples/python_native/boids_spatial3D_wrapped/boids_spatial3D.py)
predprey  : spatial-2D predator-prey with birth/death, [0,1000)^2, r = 5.
"""
import numpy as np


def circles(n=16384, seed=12, radius=2.0, repulse=0.05):
    width = np.float32(np.cbrt(np.float64(n)))
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.0, float(width), size=(3, n)).astype(np.float32)
    pos = np.minimum(pos, np.nextafter(width, np.float32(0)))
    params = np.array([width, radius, repulse], dtype=np.float32)
    drift = np.zeros(n, dtype=np.float32)
    return params, [pos[0].copy(), pos[1].copy(), pos[2].copy(), drift]


BOIDS_PARAM_NAMES = ["MIN_POSITION", "MAX_POSITION", "INTERACTION_RADIUS", "SEPARATION_RADIUS",
                     "TIME_SCALE", "GLOBAL_SCALE", "STEER_SCALE", "COLLISION_SCALE", "MATCH_SCALE",
                     "MIN_SPEED", "MAX_SPEED"]


def boids(n=1 << 20, seed=7, width=3.2, radius=0.05, speed=0.05):
    rng = np.random.default_rng(seed)
    w = np.float32(width)
    pos = rng.uniform(0.0, float(w), size=(3, n)).astype(np.float32)
    pos = np.minimum(pos, np.nextafter(w, np.float32(0)))
    vel = rng.uniform(-1.0, 1.0, size=(3, n)).astype(np.float32)
    norm = np.sqrt((vel.astype(np.float64) ** 2).sum(axis=0))
    norm[norm == 0] = 1.0
    vel = (vel / norm * speed).astype(np.float32)
    params = np.array([0.0, w, radius, 0.015, 1.0, 0.15, 0.005, 0.01, 0.125, speed * 0.5, speed],
                      dtype=np.float32)
    return params, [pos[0].copy(), pos[1].copy(), pos[2].copy(), vel[0].copy(), vel[1].copy(), vel[2].copy()]


PREDPREY_PARAM_NAMES = ["WIDTH", "RADIUS", "EAT_RADIUS", "P_PREY_DEATH", "P_PREY_REPRO", "P_PRED_REPRO",
                        "STARVE_STEPS", "JITTER", "FLEE_SCALE", "HUNT_SCALE"]


def predprey(n_prey=1_000_000, n_pred=250_000, seed=3, width=1000.0, radius=5.0):
    rng = np.random.default_rng(seed)
    w = np.float32(width)

    def pop(n, extra_u32):
        pos = rng.uniform(0.0, float(w), size=(2, n)).astype(np.float32)
        pos = np.minimum(pos, np.nextafter(w, np.float32(0)))
        vel = rng.uniform(-1.0, 1.0, size=(2, n)).astype(np.float32)
        arrs = [pos[0].copy(), pos[1].copy(), vel[0].copy(), vel[1].copy()]
        if extra_u32:
            arrs.append(np.zeros(n, dtype=np.uint32))
        return arrs

    params = np.array([w, radius, 1.0, 0.005, 0.02, 0.01, 50.0, 1.0, 0.5, 0.5], dtype=np.float32)
    return params, pop(n_prey, False), pop(n_pred, True)
