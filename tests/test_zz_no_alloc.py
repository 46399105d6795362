"""qhg_env_tables_noalloc (capacities and evaluator weights from device
arrays, with the evaluator scratch in a fixed device array) against the
population path that computes the same tables (qhg_gpu_set_env): bit-exact,
on two calls in a row, the second on a smaller grid; and the size limit."""
import ctypes as C

import numpy as np
import pytest

from qhg4_amd import abi, scenarios
from qhg4_amd._lib import lib
from qhg4_amd.population import GpuPopulation

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Dev:
    """Device buffers of the test (the caller's inputs and outputs)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.bufs = []

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(arr.nbytes, 8))) == 0
        assert self.hip.hipMemcpy(p, arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes), H2D) == 0
        self.bufs.append(p)
        return p

    def empty(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        self.bufs.append(p)
        return p

    def get(self, p, n, dtype):
        out = np.empty(n, dtype=dtype)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), D2H) == 0
        return out

    def close(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def world(seed, level, mode, cumulate):
    """A grid with altitude, NPP, coordinates, climate, coastal flags, water
    and some ice, and the C2 model's capacity and evaluator parameters."""
    sc = scenarios.c2(seed=seed, level=level, n_agents=10, patch_cells=10)
    grid, env, p = sc["grid"], sc["env"], sc["params"]
    n = grid["num_cells"]
    rng = np.random.default_rng(seed)
    env["temperature"] = rng.uniform(-5.0, 30.0, n)
    env["rainfall"] = rng.uniform(0.0, 3000.0, n)
    env["coastal"] = (rng.random(n) < 0.1).astype(np.uint8)
    env["water"] = rng.uniform(0.0, 0.3, n)
    env["ice"] = (rng.random(n) < 0.05).astype(np.uint8)
    lon, lat = env["longitude"], env["latitude"]
    oce = (lon > 115.0) & (lon < 150.0) & (lat > -12.0) & (lat < 1.0)
    env["base_npp"] = np.where(oce & (rng.random(n) < 0.5), 0.01, env["base_npp"])
    p.npp_min = 0.05
    p.npp_coastal_factor, p.npp_coastal_min_lat, p.npp_coastal_max_lat = 0.1, -40.0, 40.0
    p.npp_water_factor = 0.05
    p.me_mode = mode
    p.me_cumulate[1] = int(cumulate)
    p.me_weight[0], p.me_weight[1] = 1.5, 0.75
    return grid, env, p


def population_tables(grid, env, p):
    pop = GpuPopulation(grid, env, p)
    pop.set_env(env)
    k, w = pop.capacities(), pop.env_weights()
    pop.close()
    return k, w


def noalloc_tables(dev, grid, env, p):
    n, conn = grid["num_cells"], grid["connectivity"]
    g = abi.GridDesc()
    g.num_cells, g.connectivity = n, conn
    g.neighbors = C.cast(dev.put(grid["neighbors"].astype(np.int32)), C.POINTER(C.c_int32))
    g.num_neighbors = C.cast(dev.put(grid["num_neighbors"].astype(np.uint8)), C.POINTER(C.c_uint8))
    e = abi.EnvDesc()
    for k in ("altitude", "longitude", "latitude", "water", "base_npp", "temperature", "rainfall"):
        setattr(e, k, C.cast(dev.put(env[k].astype(np.float64)), C.POINTER(C.c_double)))
    for k in ("ice", "coastal"):
        setattr(e, k, C.cast(dev.put(env[k].astype(np.uint8)), C.POINTER(C.c_uint8)))
    K = dev.empty(n * 8)
    W = dev.empty(n * (conn + 1) * 8)
    rc = lib().qhg_env_tables_noalloc(C.byref(g), C.byref(e), C.byref(p), K, W, None)
    assert rc == 0, lib().qhg_gpu_last_error().decode()
    return dev.get(K, n, np.float64), dev.get(W, n * (conn + 1), np.float64)


def _bits(x):
    return np.ascontiguousarray(x).view(np.uint64)


def test_noalloc_tables_equal_the_population_path():
    dev = Dev()
    try:
        # two calls in a row on one GPU, the second on a smaller grid
        for seed, level, mode, cum in ((61, 6, abi.ME_MUL_SIMPLE, False), (62, 4, abi.ME_ADD_SIMPLE, True)):
            print("case: qhg_env_tables_noalloc, ico level %d, mode %d, cumulate %d" % (level, mode, cum),
                  flush=True)
            grid, env, p = world(seed, level, mode, cum)
            k_ref, w_ref = population_tables(grid, env, p)
            k, w = noalloc_tables(dev, grid, env, p)
            assert np.array_equal(_bits(k), _bits(k_ref)), "capacities differ at level %d" % level
            assert np.array_equal(_bits(w), _bits(w_ref)), "weights differ at level %d" % level
            assert np.count_nonzero(k) > 0 and np.count_nonzero(w) > 0
    finally:
        dev.close()


def test_noalloc_refuses_grids_beyond_its_scratch():
    g = abi.GridDesc()
    g.num_cells, g.connectivity = (1 << 20) + 1, 6
    e = abi.EnvDesc()
    p = scenarios.ooa_params(1)
    rc = lib().qhg_env_tables_noalloc(C.byref(g), C.byref(e), C.byref(p), None, None, None)
    assert rc == -3
    assert "scratch" in lib().qhg_gpu_last_error().decode()
