"""Serving: the request loop of ``stater_server_*`` and a Poisson load driver.

``Server`` binds the C-ABI serving runtime (admission by device-memory
budget, one worker and HIP stream per slot, a job's statistics resident
across its adaptive-sampling rounds).  ``run_poisson`` submits jobs of a
resolution / transform-count / sample-count mix at exponential inter-arrival
times and reports throughput and latency percentiles.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from . import (MAX_LAMBDAS, STATER_ACC_SEQUENTIAL, STATER_BF16, STATER_F32, Context, GBuffers, Job, JobResult,
               Plane, Samples, ServerConfig, check)
from .workload import FEATURE_SDS, MT5_LAMBDAS

NONID = tuple(l for l in MT5_LAMBDAS if l != 1.0)


def job_lambdas(T: int) -> tuple[float, ...]:
    """The first T non-identity lambdas of the mt5 list plus the identity."""
    return NONID[:T] + (1.0,)


class Server:
    def __init__(self, ctx: Context, max_bytes: int = 64 << 30, slots: int = 4, chunk_spp: int = 32,
                 max_width: int = 3840, max_height: int = 2160, acc_mode: int = STATER_ACC_SEQUENTIAL,
                 params=None, feature_sds=FEATURE_SDS):
        self.ctx, self.lib = ctx, ctx.lib
        cfg = ServerConfig()
        self.lib.stater_server_default_config(C.byref(cfg))
        cfg.max_bytes, cfg.slots, cfg.chunk_spp = max_bytes, slots, chunk_spp
        cfg.max_width, cfg.max_height, cfg.acc_mode = max_width, max_height, acc_mode
        for k in range(3):
            cfg.feature_sds[k] = feature_sds[k] if k < len(feature_sds) else 0.0
        if params is not None:
            cfg.params = params
        self.cfg = cfg
        h = C.c_void_p()
        check(self.lib.stater_server_create(ctx.h, C.byref(cfg), C.byref(h)))
        self.h = h
        self._refs = {}  # caller structs a submitted request points to, until it is waited for
        ctx._users.add(self)

    @staticmethod
    def job(width: int, height: int, lambdas=MT5_LAMBDAS, spp: int = 64, rounds: int = 1, cap: int = 4096,
            model: int = 2, seed: int = 5, image_id: int = 0, bf16: bool = False, keep: bool = False,
            film: bool = False, samples: Samples | None = None, gbufs: GBuffers | None = None,
            continue_id: int = 0, next_spp: int = 0) -> Job:
        """A job: samples from the seeded scene (model, seed, image_id), or the
        caller's chunk `samples` with features `gbufs`; continue_id continues
        a kept job; next_spp > 0 computes the AS map of a next round."""
        j = Job()
        j.width, j.height, j.n_lambdas = width, height, len(lambdas)
        assert len(lambdas) <= MAX_LAMBDAS
        for i, l in enumerate(lambdas):
            j.lambdas[i] = l
        j.spp, j.rounds, j.cap_spp, j.scene_model = spp, rounds, cap, model
        j.seed, j.image_id = seed, image_id
        j.feature_dtype = STATER_BF16 if bf16 else STATER_F32
        j.keep = 1 if keep else 0
        j.film = 1 if film else 0
        j.next_spp = next_spp
        j.continue_id = continue_id
        j._refs = (samples, gbufs)
        j.samples = C.addressof(samples) if samples is not None else None
        j.gbufs = C.addressof(gbufs) if gbufs is not None else None
        return j

    def job_bytes(self, job: Job) -> int:
        b = self.lib.stater_server_job_bytes(self.h, C.byref(job))
        if b < 0:
            check(int(b))
        return int(b)

    def submit(self, job: Job) -> int:
        jid = C.c_uint64()
        check(self.lib.stater_server_submit(self.h, C.byref(job), C.byref(jid)))
        self._refs[jid.value] = getattr(job, "_refs", None)
        return jid.value

    def wait(self, jid: int) -> dict:
        r = JobResult()
        try:
            check(self.lib.stater_server_wait(self.h, jid, C.byref(r)))
        finally:
            self._refs.pop(jid, None)
        return {"id": r.id, "status": r.status, "slot": r.slot, "t_submit": r.t_submit, "t_admit": r.t_admit,
                "t_done": r.t_done, "samples": r.samples, "bytes": r.bytes}

    def view(self, jid: int, name: str) -> Plane:
        p = Plane()
        check(self.lib.stater_server_view(self.h, jid, name.encode(), C.byref(p)))
        return p

    def release(self, jid: int) -> None:
        check(self.lib.stater_server_release(self.h, jid))

    def drain(self) -> None:
        check(self.lib.stater_server_drain(self.h))

    def clock(self) -> float:
        return self.lib.stater_server_clock(self.h)

    def close(self) -> None:
        if self.h:
            self.lib.stater_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Mix:
    """Job mix: resolutions with weights, transform counts T, spp values."""
    sizes: tuple = ((1280, 720, 0.50), (1920, 1080, 0.35), (3840, 2160, 0.15))
    ts: tuple = (1, 3, 5)
    spps: tuple = (64, 256, 1024)
    rounds: int = 1
    bf16: bool = False
    film: bool = True  # the renderer's loop: guarded samples, pbrt Film into the Welch filter

    def draw(self, rng: np.random.Generator, image_id: int, seed: int) -> Job:
        w = np.array([s[2] for s in self.sizes])
        W, H, _ = self.sizes[rng.choice(len(self.sizes), p=w / w.sum())]
        T = int(rng.choice(self.ts))
        spp = int(rng.choice(self.spps))
        return Server.job(W, H, job_lambdas(T), spp=spp, rounds=self.rounds, seed=seed, image_id=image_id,
                          bf16=self.bf16, film=self.film)


def run_poisson(server: Server, rate: float, n_jobs: int, mix: Mix, seed: int = 5,
                log=None) -> dict:
    """Submits n_jobs at Poisson arrivals of `rate` jobs/s (arrival times on the
    server clock); returns throughput and latency percentiles."""
    rng = np.random.default_rng(seed)
    jobs = [mix.draw(rng, i, 5000 + i) for i in range(n_jobs)]
    gaps = rng.exponential(1.0 / rate, n_jobs)
    t0 = server.clock()
    arrivals = t0 + np.cumsum(gaps)
    ids = []
    for k, (job, ta) in enumerate(zip(jobs, arrivals)):
        dt = ta - server.clock()
        if dt > 0:
            time.sleep(dt)
        ids.append(server.submit(job))
        if log and (k + 1) % max(1, n_jobs // 5) == 0:
            log(f"submitted {k + 1}/{n_jobs}")
    res = [server.wait(i) for i in ids]
    lat = np.array([r["t_done"] - r["t_submit"] for r in res])
    queue = np.array([r["t_admit"] - r["t_submit"] for r in res])
    t_end = max(r["t_done"] for r in res)
    px = sum(j.width * j.height for j in jobs)
    samples = sum(r["samples"] for r in res)
    span = t_end - t0
    return {
        "jobs": n_jobs,
        "rate_jobs_s": rate,
        "mpixel_s": px / span / 1e6,
        "gsample_s": samples / span / 1e9,
        "jobs_s": n_jobs / span,
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "queue_p50_ms": float(np.percentile(queue, 50) * 1e3),
        "span_s": span,
        "mpixel": px / 1e6,
    }
