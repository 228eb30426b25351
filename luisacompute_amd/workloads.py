"""Synthetic workloads of the benchmark configurations, built from the PCG32 generators of
lc-backend-hip.so (stream initstate 42, initseq = workload number).

  icosphere  : level-6 icosphere (81,920 triangles), 1,048,576 rays from r = 3 toward the unit ball
  soup       : 16,777,216-triangle soup in [0,100]^3 (sigma 0.5), 8,388,608 rays in the cube
  instanced  : 1,024 instances of 8 meshes (icospheres L3..L8, soups of 200k and 500k triangles in
               [-1,1]^3 with sigma 0.01, i.e. the soup recipe scaled by 1/50), rotation uniform
               (Shoemake), translation U[-50,50]^3, scale U[0.5,2]; 64 ray batches with sizes
               log-uniform in [4096, 4194304]; rays: origin U[-60,60]^3, direction U(S^2).
Stream layout of the instanced workload (initseq 4): 8 draws per instance, then one draw per batch
size, then 5 draws per ray, batch after batch. Soup meshes use initseq 40 and 41.
"""
import numpy as np

from . import icosphere, soup, instances, batch_sizes, rays_box_sphere

INSTANCED_MESHES_FULL = (("ico", 3), ("ico", 4), ("ico", 5), ("ico", 6), ("ico", 7), ("ico", 8),
                         ("soup", 200000), ("soup", 500000))
INSTANCED_MESHES_SMALL = (("ico", 1), ("ico", 2), ("ico", 3), ("ico", 3), ("ico", 4), ("ico", 4),
                          ("soup", 2000), ("soup", 5000))


def mesh_arrays(kind, arg, seq):
    if kind == "ico":
        return icosphere(arg)
    v, t = soup(arg, initstate=42, initseq=seq, extent=2.0, sigma=0.01)
    v[:, :3] -= np.float32(1.0)  # centre the soup on the origin
    return v, t


class Instanced:
    """The instanced workload's arrays (meshes, instance table, batches)."""

    def __init__(self, small=False, instances_count=1024, batches=64, lo=None, hi=None):
        spec = INSTANCED_MESHES_SMALL if small else INSTANCED_MESHES_FULL
        lo = lo or (16 if small else 4096)
        hi = hi or (4096 if small else 4194304)
        self.meshes = [mesh_arrays(k, a, 40 + i % 2 if k == "soup" else 0) for i, (k, a) in enumerate(spec)]
        self.affines, self.mesh_ids = instances(instances_count, len(spec), initstate=42, initseq=4)
        skip = 8 * instances_count
        self.batch_sizes = batch_sizes(batches, lo, hi, initstate=42, initseq=4, skip=skip)
        self.ray_skip0 = skip + batches
        self.extent = 60.0

    def batch_rays(self, j):
        skip = self.ray_skip0 + 5 * int(np.sum(self.batch_sizes[:j], dtype=np.uint64))
        return rays_box_sphere(int(self.batch_sizes[j]), initstate=42, initseq=4, skip=skip, lo=-self.extent,
                               hi=self.extent)

    @property
    def total_rays(self):
        return int(np.sum(self.batch_sizes, dtype=np.uint64))
