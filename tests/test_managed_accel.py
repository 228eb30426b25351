"""ManagedAccel bookkeeping (src/py/managed_accel.cpp:22-157) on recording stand-ins of the
device objects: instances with the same geometry description share one BLAS, new descriptions are
built once at the next update, a BLAS is released with its last instance."""
import pytest

from luisacompute_amd import LCError
from luisacompute_amd.managed import ManagedAccel


class _Cmd:
    def __init__(self, what, obj):
        self.what, self.obj = what, obj


class _Blas:
    def __init__(self, log, kind, args):
        self.log, self.kind, self.args = log, kind, args

    def build(self, request=0):
        return _Cmd("build " + self.kind, self)


class _Accel:
    def __init__(self):
        self.slots = []

    def emplace_back(self, blas, *a):
        self.slots.append(blas)

    def set(self, i, blas, *a):
        self.slots[i] = blas

    def pop_back(self):
        self.slots.pop()

    def build(self, request=0):
        return _Cmd("build accel", self)

    def update_instance_buffer(self):
        return _Cmd("instances", self)


class _Device:
    def __init__(self):
        self.created = []

    def create_accel(self, option=None):
        return _Accel()

    def create_mesh(self, vb, tb, **kw):
        b = _Blas(self.created, "mesh", (vb, tb, kw))
        self.created.append(b)
        return b

    def create_procedural_primitive(self, ab, **kw):
        b = _Blas(self.created, "procedural", (ab, kw))
        self.created.append(b)
        return b


class _Stream:
    def __init__(self):
        self.cmds = []

    def __lshift__(self, c):
        self.cmds.append(c)
        return self

    def synchronize(self):
        pass


class _Buf:
    def __init__(self, stride, count):
        self.stride, self.count, self.nbytes = stride, count, stride * count


def test_sharing_building_and_release():
    dev, s = _Device(), _Stream()
    m = ManagedAccel(dev)
    vb, tb, vb2 = _Buf(16, 30), _Buf(12, 10), _Buf(16, 30)
    m.add(vb, tb)
    m.add(vb, tb)                      # same description: shared
    assert m.blas_count == 1 and len(dev.created) == 1 and len(m) == 2
    assert dev.created[0].args[0] is vb and dev.created[0].args[1] is tb
    m.update(s)
    assert [c.what for c in s.cmds] == ["build mesh", "build accel"]
    s.cmds.clear()
    m.update(s)                        # nothing new: only the TLAS
    assert [c.what for c in s.cmds] == ["build accel"]
    m.set(0, vb2, tb)                  # new description for instance 0
    assert m.blas_count == 2
    s.cmds.clear()
    m.update(s)
    assert [c.what for c in s.cmds] == ["build mesh", "build accel"] and s.cmds[0].obj is dev.created[1]
    m.set(1, vb2, tb)                  # the first BLAS loses its last instance
    assert m.blas_count == 1 and m.accel.slots == [dev.created[1]] * 2
    m.add_buffer_view(vb, 16, 160, 16, tb, 0, 120)   # a sub-range is another description
    assert m.blas_count == 2
    m.pop()
    assert m.blas_count == 1 and len(m) == 2


def test_procedural_descriptions_and_checks():
    dev, s = _Device(), _Stream()
    m = ManagedAccel(dev)
    ab = _Buf(24, 100)
    m.add_procedural(ab)
    m.add_procedural(ab, 0, 100)       # the same range
    m.add_procedural(ab, 10, 20)
    assert m.blas_count == 2
    assert dev.created[1].args[0] is ab
    assert dev.created[1].args[1]["offset"] == 240 and dev.created[1].args[1]["size"] == 480
    with pytest.raises(LCError):
        m.add_procedural(_Buf(16, 10))
    with pytest.raises(LCError):
        m.add_buffer_view(_Buf(16, 4), 32, 64, 16, _Buf(12, 1), 0, 12)
    # a description created and dropped before any update is never built
    m.set_procedural(2, ab, 0, 5)
    m.set_procedural(2, ab)
    m.update(s)
    assert [c.what for c in s.cmds] == ["build procedural", "build accel"]
