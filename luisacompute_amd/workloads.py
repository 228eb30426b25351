"""Synthetic workloads of the benchmark configurations, built from the PCG32 generators of
libluisa_compute_backend_hip.so (stream initstate 42, initseq = workload number).

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


# ---------------------------------------------------------------------------------------------
# Serving workload: four resident scenes and a Poisson request stream.
RESOLUTIONS = (256, 512, 1024, 2048)
SPPS = (1, 4, 16, 64)


def serving_requests(n, rate, initstate=42, initseq=5):
    """n requests (PCG32 stream (initstate, initseq), 4 draws each): scene floor(4u), square
    resolution from RESOLUTIONS, spp from SPPS, Poisson arrivals at `rate` per second
    (exponential gaps -ln(1 - u) / rate). Returns a list of dicts with the arrival time."""
    u = pcg32_floats(initstate, initseq, 4 * n).astype(np.float64).reshape(n, 4)
    t = 0.0
    out = []
    for k in range(n):
        t += -np.log1p(-u[k, 3]) / rate if rate > 0 else 0.0
        out.append(dict(index=k, scene=int(u[k, 0] * 4), res=RESOLUTIONS[int(u[k, 1] * 4)], spp=SPPS[int(u[k, 2] * 4)],
                        arrival=t, sequence=(5 << 32) + (k << 22)))
    return out


def pt_scene(light_position, light_u, light_v, emission, origin, right, up, forward):
    from . import PtScene
    s = PtScene()
    for name, val in (("light_position", light_position), ("light_u", light_u), ("light_v", light_v),
                      ("light_emission", emission), ("camera_origin", origin), ("camera_right", right),
                      ("camera_up", up), ("camera_forward", forward)):
        getattr(s, name)[:] = [float(np.float32(x)) for x in val]
    return s


def serving_scenes(device, stream, small=False):
    """The four resident scenes of the serving workload: (name, accel, pt_instances, PtScene, keep).
    icosphere L6 (C1), the 16,777,216-triangle soup (C2), the box room (C3) and the 1,024-instance
    TLAS (C4); each lit by a quad light above it and viewed from outside (the room from inside).
    small=True scales every scene down for tests."""
    from . import PtInstance, host
    scenes = []
    # 1. icosphere
    v, t = icosphere(3 if small else 6)
    acc = host.single_mesh_accel(device, stream, v, t)
    inst = [PtInstance()]
    inst[0].albedo[:] = [0.7, 0.5, 0.3]
    scenes.append(("icosphere", acc, inst, pt_scene((-1, 3, -1), (2, 0, 0), (0, 0, 2), (10, 10, 10),
                                                    (0, 0, 3.5), (0.5, 0, 0), (0, 0.5, 0), (0, 0, -1))))
    # 2. soup
    ntri = 20000 if small else 1 << 24
    v, t = soup(ntri, initstate=42, initseq=2)
    acc = host.single_mesh_accel(device, stream, v, t)
    inst = [PtInstance()]
    inst[0].albedo[:] = [0.5, 0.5, 0.5]
    scenes.append(("soup", acc, inst, pt_scene((20, 150, 20), (60, 0, 0), (0, 0, 60), (20, 20, 20),
                                               (50, 50, -60), (-0.6, 0, 0), (0, 0.6, 0), (0, 0, 1))))
    # 3. box room
    room = BoxRoom(wall_quads=16, block_quads=4) if small else BoxRoom()
    acc, _, inst = room.build(device, stream)
    scenes.append(("box room", acc, inst, room.scene()))
    # 4. instanced
    w = Instanced(small=small)
    gm = [host.mesh_from_arrays(device, stream, mv, mt) for mv, mt in w.meshes]
    acc = device.create_accel()
    for aff, k in zip(w.affines, w.mesh_ids):
        acc.emplace_back(gm[int(k)], aff)
    stream << acc.build()
    acc._keep = gm
    alb = 0.1 + 0.8 * pcg32_floats(42, 6, 3 * len(w.mesh_ids)).astype(np.float64).reshape(-1, 3)
    inst = []
    for k in range(len(w.mesh_ids)):
        p = PtInstance()
        p.albedo[:] = [float(x) for x in alb[k]]
        inst.append(p)
    scenes.append(("instanced", acc, inst, pt_scene((-50, 120, -50), (100, 0, 0), (0, 0, 100), (5, 5, 5),
                                                    (0, 0, -140), (-0.5, 0, 0), (0, 0.5, 0), (0, 0, 1))))
    return scenes


# ---------------------------------------------------------------------------------------------
# Box room of the path-tracing workload.
def pcg32_floats(initstate, initseq, n):
    """n PCG32 floats (x >> 8) * 2^-24 of stream (initstate, initseq) (O'Neill 2014)."""
    M = (1 << 64) - 1
    mult = 6364136223846793005
    inc = ((initseq << 1) | 1) & M
    state = 0

    def step(s):
        old = s
        s = (old * mult + inc) & M
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return s, ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    state, _ = step(state)
    state = (state + initstate) & M
    state, _ = step(state)
    out = []
    for _ in range(n):
        state, x = step(state)
        out.append(np.float32((x >> 8) * 2.0 ** -24))
    return np.array(out, np.float32)


def grid_quad(origin, a, b, n):
    """n x n quads spanning origin + [0,1] a + [0,1] b, two triangles each, facing cross(a, b)."""
    origin, a, b = (np.asarray(x, np.float64) for x in (origin, a, b))
    s = np.arange(n + 1, dtype=np.float64) / n
    I, J = np.meshgrid(s, s, indexing="ij")
    p = origin + I[..., None] * a + J[..., None] * b
    verts = np.zeros(((n + 1) * (n + 1), 4), np.float32)
    verts[:, :3] = p.reshape(-1, 3)
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    v00 = (i * (n + 1) + j).ravel()
    v10 = ((i + 1) * (n + 1) + j).ravel()
    v11 = ((i + 1) * (n + 1) + j + 1).ravel()
    v01 = (i * (n + 1) + j + 1).ravel()
    tris = np.empty((2 * n * n, 3), np.uint32)
    tris[0::2] = np.stack([v00, v10, v11], 1)
    tris[1::2] = np.stack([v00, v11, v01], 1)
    return verts, tris


def merge(parts):
    verts, tris, base = [], [], 0
    for v, t in parts:
        verts.append(v)
        tris.append(t + base)
        base += v.shape[0]
    return np.concatenate(verts), np.concatenate(tris).astype(np.uint32)


def block(lo, hi, m):
    """Axis-aligned box [lo, hi], m x m quads per face, faces pointing outward."""
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    e = hi - lo
    X, Y, Z = np.array([e[0], 0, 0]), np.array([0, e[1], 0]), np.array([0, 0, e[2]])
    faces = [((hi[0], lo[1], lo[2]), Y, Z), (lo, Z, Y),                 # +x, -x
             ((lo[0], hi[1], lo[2]), Z, X), (lo, X, Z),                 # +y, -y
             ((lo[0], lo[1], hi[2]), X, Y), (lo, Y, X)]                 # +z, -z
    return merge([grid_quad(o, a, b, m) for o, a, b in faces])


class BoxRoom:
    """The path-tracing workload's scene: a closed room [-1,1] x [0,2] x [-1,1] of 6 walls, each
    wall_quads^2 quads (2 triangles each) facing inward; 2 axis-aligned blocks on the floor, each
    face block_quads^2 quads; a quad light of radiance `emission` just below the ceiling, facing
    down. Albedos: U[0.1, 0.9] per channel for the 6 walls and 2 blocks, in that order, from PCG32
    stream (42, 3). Instances (identity transforms): floor, ceiling, back, front, left, right,
    short block, tall block, light. The camera sits inside the room by the front wall, looking
    along -z with a vertical half-angle of 45 degrees (the light is in view)."""

    LIGHT_POSITION = (-0.24, 1.98, -0.22)
    LIGHT_U = (0.47, 0.0, 0.0)
    LIGHT_V = (0.0, 0.0, 0.38)

    def __init__(self, wall_quads=512, block_quads=128, emission=15.0, initstate=42, initseq=3, albedo=None):
        n = wall_quads
        walls = [((-1, 0, -1), (0, 0, 2), (2, 0, 0)),   # floor, facing +y
                 ((-1, 2, -1), (2, 0, 0), (0, 0, 2)),   # ceiling, -y
                 ((-1, 0, -1), (2, 0, 0), (0, 2, 0)),   # back, +z
                 ((-1, 0, 1), (0, 2, 0), (2, 0, 0)),    # front, -z
                 ((-1, 0, -1), (0, 2, 0), (0, 0, 2)),   # left, +x
                 ((1, 0, -1), (0, 0, 2), (0, 2, 0))]    # right, -x
        self.meshes = [grid_quad(o, a, b, n) for o, a, b in walls]
        self.meshes.append(block((0.05, 0.0, -0.05), (0.65, 0.6, 0.55), block_quads))    # short block
        self.meshes.append(block((-0.65, 0.0, -0.65), (-0.05, 1.2, -0.05), block_quads))  # tall block
        p, u, v = (np.asarray(x, np.float64) for x in (self.LIGHT_POSITION, self.LIGHT_U, self.LIGHT_V))
        lv = np.zeros((4, 4), np.float32)
        lv[:, :3] = [p, p + u, p + v, p + u + v]
        self.meshes.append((lv, np.array([[0, 1, 2], [1, 3, 2]], np.uint32)))
        if albedo is None:
            albedo = 0.1 + 0.8 * pcg32_floats(initstate, initseq, 24).astype(np.float64).reshape(8, 3)
        self.albedo = np.zeros((9, 3), np.float32)
        self.albedo[:8] = np.asarray(albedo, np.float32).reshape(8, 3)
        self.emissive = np.array([0] * 8 + [1], np.uint32)
        self.emission = (emission, emission, emission)
        t = 1.0
        self.camera = dict(origin=(0.0, 1.0, 0.98), right=(t, 0.0, 0.0), up=(0.0, t, 0.0), forward=(0.0, 0.0, -1.0))

    @property
    def triangle_count(self):
        return sum(t.shape[0] for _, t in self.meshes)

    def scene(self):
        from . import PtScene
        s = PtScene()
        for name, val in (("light_position", self.LIGHT_POSITION), ("light_u", self.LIGHT_U),
                          ("light_v", self.LIGHT_V), ("light_emission", self.emission),
                          ("camera_origin", self.camera["origin"]), ("camera_right", self.camera["right"]),
                          ("camera_up", self.camera["up"]), ("camera_forward", self.camera["forward"])):
            getattr(s, name)[:] = [float(np.float32(x)) for x in val]
        return s

    def build(self, device, stream):
        """Uploads the meshes, builds them and an accel of identity instances; returns
        (accel, meshes, pt_instances)."""
        from . import PtInstance, host
        meshes = [host.mesh_from_arrays(device, stream, v, t) for v, t in self.meshes]
        accel = device.create_accel()
        for m in meshes:
            accel.emplace_back(m)
        stream << accel.build()
        inst = []
        for k, m in enumerate(meshes):
            pi = PtInstance()
            pi.albedo[:] = [float(x) for x in self.albedo[k]]
            pi.emissive = int(self.emissive[k])
            inst.append(pi)
        accel._keep = meshes
        return accel, meshes, inst
