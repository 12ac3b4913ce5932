"""``turbomodal.io.load_mesh`` (python/turbomodal/io.py:1955-1990) for gmsh MSH 2.x sector
meshes: the native loader, then the cyclic boundaries identified and matched.  The
reference's meshio formats (NASTRAN, Abaqus, VTK, CGNS, MED, XDMF) and CAD import are
not part of this package and raise."""
from __future__ import annotations

from pathlib import Path

from turbomodal_amd import Mesh


def load_mesh(filepath, num_sectors: int, file_format: str | None = None) -> Mesh:
    filepath = Path(filepath)
    if not filepath.exists():
        raise FileNotFoundError(f"Mesh file not found: {filepath}")
    if filepath.suffix.lower() != ".msh" or file_format is not None:
        raise ValueError(f"Unsupported mesh format '{filepath.suffix.lower()}': this package reads gmsh MSH 2.x "
                         "(.msh) sector meshes only")
    mesh = Mesh()
    mesh.num_sectors = num_sectors
    mesh.load_from_gmsh(str(filepath))
    mesh.identify_cyclic_boundaries()
    mesh.match_boundary_nodes()
    return mesh
