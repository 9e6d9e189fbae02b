"""sofa_amd: SOFA's implicit corotational-FEM time step (TetrahedronFEMForceField "large" +
CGLinearSolver + EulerImplicitSolver) on AMD MI355X (gfx950).

The compute path is libsofa_amd.so (hand-written HIP, C ABI in include/sofa_amd.h); this
package binds it (``capi``, which loads the library on first use and raises if it is
missing), builds scene inputs (``meshes``) and mirrors SOFA's scene-graph interface for the
components on the path (``scene``).
"""
from . import meshes  # noqa: F401
from . import capi  # noqa: F401

__all__ = ["meshes", "capi"]
