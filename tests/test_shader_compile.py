"""Run-time shaders compile for gfx950 against the embedded device library (hiprtc works without
a GPU), and compile errors come back with the compiler's diagnostics."""
import pytest

import luisacompute_amd as lc
import shaders


@pytest.mark.parametrize("entry", sorted(shaders.ALL))
def test_shader_sources_compile(entry):
    assert lc.compile_shader_source(shaders.ALL[entry]) > 1000


def test_compile_errors_carry_the_log():
    bad = shaders.TRACE.replace("dev::trace_closest(", "dev::trace_closestt(")
    with pytest.raises(lc.LCError, match="trace_closestt"):
        lc.compile_shader_source(bad)
