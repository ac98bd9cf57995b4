"""Source rules for the HIP kernels that a compiler run does not enforce.

Ordered device-scope atomics and fences: on MI355X an acquire or release
at agent scope writes back and invalidates the XCD's L2, which made a
reduction whose every block took an acquire-release counter read at 4.8
instead of 6.0 TB/s.  Kernels communicate through relaxed agent-scope
atomics on words that carry their own flag; the only ordered atomic is the
system-scope release that posts a scalar to the host once per call.

hipcc (ROCm 7.2, clang 22) lowers __builtin_bit_cast applied directly to an
ext_vector subscript, e.g. __builtin_bit_cast(int, w[1]) with w an
ext_vector_type(4) value, as a read of element 0 whatever the index: the
kernels then silently use the wrong element.  Such casts go through a scalar
first (partition.hip's elem_as); this test keeps the direct form out.
"""
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CAST_OF_SUBSCRIPT = re.compile(r"__builtin_bit_cast\(\s*[^,()]+,\s*[A-Za-z_][A-Za-z_0-9.]*\s*\[")


def test_no_bit_cast_of_vector_subscript():
    files = glob.glob(os.path.join(ROOT, "neslgpu_amd", "csrc", "*.hip")) + \
        glob.glob(os.path.join(ROOT, "neslgpu_amd", "csrc", "*.hpp"))
    assert files
    bad = []
    for f in files:
        for i, line in enumerate(open(f), 1):
            if CAST_OF_SUBSCRIPT.search(line):
                bad.append(f"{os.path.basename(f)}:{i}: {line.strip()}")
    assert not bad, "bit_cast of a subscript (cast a scalar copy instead):\n" + "\n".join(bad)


def test_rule_catches_the_pattern():
    assert CAST_OF_SUBSCRIPT.search("v[0] = __builtin_bit_cast(T, a[1]);")
    assert not CAST_OF_SUBSCRIPT.search("x[c] = elem_as<T>(w[c]);")
    assert not CAST_OF_SUBSCRIPT.search("__builtin_bit_cast(double, ((uint64_t)w.y << 32) | w.x)")


ORDERED = re.compile(r"__threadfence\s*\(|"
                     r"__builtin_amdgcn_fence\s*\(\s*__ATOMIC_\w+\s*,\s*\"(agent)?\"|"
                     r"__hip_atomic_\w+\s*\([^;]*__ATOMIC_(ACQ_REL|SEQ_CST|ACQUIRE|RELEASE)[^;]*__HIP_MEMORY_SCOPE_AGENT")


def test_no_ordered_device_atomics_in_kernels():
    files = glob.glob(os.path.join(ROOT, "neslgpu_amd", "csrc", "*.hip")) + \
        glob.glob(os.path.join(ROOT, "neslgpu_amd", "csrc", "*.hpp"))
    bad = []
    for f in files:
        for i, line in enumerate(open(f), 1):
            if ORDERED.search(line):
                bad.append(f"{os.path.basename(f)}:{i}: {line.strip()}")
    assert not bad, "ordered device-scope atomic or fence in kernel code:\n" + "\n".join(bad)


def test_ordered_rule_catches_the_patterns():
    assert ORDERED.search("prev = __hip_atomic_fetch_add(c, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);")
    assert ORDERED.search("__threadfence();")
    assert ORDERED.search('__builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");')
    assert not ORDERED.search('__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");')
    assert not ORDERED.search("__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);")
    assert not ORDERED.search("__hip_atomic_store(&mb->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);")
