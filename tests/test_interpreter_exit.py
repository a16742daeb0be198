"""Objects still alive when the interpreter exits are left to the process
end: their finalizers run after the modules' globals are cleared, and
calling into the libraries from there printed "Exception ignored ...
TypeError" for every such object."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = """
import torch
from h2lib_amd import geometry as G
from h2lib_amd.geometry import BlockTree, ClusterTree
mesh = G.sphere(4)
tree = ClusterTree(mesh, 1, 16)
blocks = BlockTree(tree, tree, 2.0, 0)
print("built", mesh.triangles)
"""


def test_exit_with_live_objects_is_quiet():
    env = dict(os.environ, PYTHONPATH=ROOT)
    r = subprocess.run([sys.executable, "-c", SCRIPT], env=env, cwd=ROOT,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "built 128" in r.stdout
    assert "Exception ignored" not in r.stderr, r.stderr
