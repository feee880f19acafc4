"""The host-side node graph (incidence, neighbourhoods and the gather slots every assembly kernel
reads) is built on several threads for large meshes; tests/cpp/test_node_graph checks that the
threaded build equals the single-threaded one array for array on hex8, hex20, tet4 and tet10 layers
(host only)."""
import os
import subprocess

EXE = os.path.join(os.path.dirname(__file__), "cpp", "test_node_graph")


def test_threaded_node_graph_equals_serial():
    assert os.path.exists(EXE), "build tests/cpp first (__graft_entry__.build())"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
