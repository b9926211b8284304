"""Grid Hamiltonian files (Hmt.xml / Hmt.json): the on-disk input of the Green-function path.

load_hamiltonian follows green_input::load_Hamiltonian (src/green_input.cxx:29-221) and its JSON branch
(include/json_reading.hxx:61-229); write_hamiltonian writes the format of grid_operator_t::write_to_file
(include/grid_operators.hxx:460-613).  The reference reads with rapidxml / rapidjson, neither of which is in
this tree; here the standard library parses.

The loaded Hamiltonian has the attributes the plan builder and the solver wrappers read (ng, bc, hg, h, dvol,
xyzZinso, natoms, V, hs), so it stands in for a synthetic system: GreenPlan(load_hamiltonian("Hmt.xml"), r_trunc).
"""
from __future__ import annotations

import json
import warnings
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field

import numpy as np


def nsho(numax: int) -> int:
    return (numax + 1) * (numax + 2) * (numax + 3) // 6 if numax >= 0 else 0


@dataclass
class Hamiltonian:
    ng: tuple                   # grid points (nx, ny, nz)
    hg: tuple                   # grid spacings (hx, hy, hz), Bohr
    bc: tuple                   # boundary conditions (0 isolated, 1 periodic, 2 vacuum, 3 repeat)
    V: np.ndarray               # (nz, ny, nx) potential, Hartree
    xyzZinso: np.ndarray        # (natoms, 8): x, y, z (centred frame), Z, global id, numax, sigma, 0
    hs: list = field(default_factory=list)   # per atom (2, nc, nc): hamiltonian then overlap, normalised SHO basis

    @property
    def natoms(self):
        return self.xyzZinso.shape[0]

    @property
    def nb(self):
        return tuple(n // 4 for n in self.ng)

    @property
    def h(self):
        """The spacing of an isotropic grid (the synthetic systems' one number)."""
        return self.hg[0]

    @property
    def dvol(self):
        return float(self.hg[0] * self.hg[1] * self.hg[2])


def _numbers(text):
    return np.array([float(t) for t in (text or "").split()], dtype=np.float64)


def _potential(values, ng, filename):
    """src/green_input.cxx:185-205: a count other than nx ny nz is zero-filled with a warning (the reference's
    default +green_input.empty.potential=1)."""
    n = int(ng[0]) * int(ng[1]) * int(ng[2])
    values = np.asarray(values, dtype=np.float64).reshape(-1)
    if values.size != n:
        warnings.warn(f"in file \"{filename}\" {ng[2]}*{ng[1]}*{ng[0]} = {n} potential values set to zero "
                      f"(found {values.size})")
        values = np.zeros(n)
    return values.reshape(int(ng[2]), int(ng[1]), int(ng[0]))


def _matrix(v, nc, name, gid, filename):
    v = np.asarray(v, dtype=np.float64).reshape(-1)
    if v.size != nc * nc:
        raise ValueError(f"{name} matrix of atom {gid} in \"{filename}\" has {v.size} values, expected {nc * nc}")
    m = v.reshape(nc, nc)
    dev = np.abs(m - m.T).max() if nc else 0.0
    if dev > 1e-6:
        warnings.warn(f"{name} matrix of atom #{gid} has a max deviation of {dev:.1e} from symmetric")
    return m


def _load_xml(filename):
    root = ET.parse(filename).getroot()
    if root.tag != "grid_Hamiltonian":
        raise ValueError(f"no grid_Hamiltonian found in file \"{filename}\"")
    ng, hg, bc = [0, 0, 0], [1.0, 1.0, 1.0], [0, 0, 0]
    xyz, hs = [], []
    atoms = root.find("sho_atoms")
    if atoms is not None:
        natoms = int(atoms.get("number", "0"))
        for atom in list(atoms):
            gid = int(atom.get("global_id", "-1"))
            Z = float(atom.get("Z", "0"))
            pos = atom.find("position")
            p = [float(pos.get(a)) if pos is not None and pos.get(a) is not None else 0.0 for a in "xyz"]
            prj = atom.find("projectors")
            numax = int(prj.get("numax", "-1")) if prj is not None else -1
            sigma = float(prj.get("sigma", "-1")) if prj is not None else -1.0
            nc = nsho(numax)
            m = np.zeros((2, nc, nc))
            for k, tag in enumerate(("hamiltonian", "overlap")):
                node = atom.find(tag)
                if node is None:
                    warnings.warn(f"atom with global_id={gid} has no {tag} matrix in file \"{filename}\"")
                    continue
                m[k] = _matrix(_numbers(node.text), nc, tag, gid, filename)
            xyz.append([p[0], p[1], p[2], Z, gid, numax, sigma, 0.0])
            hs.append(m)
        if natoms != len(xyz):
            raise ValueError(f"Inconsistency in file \"{filename}\": number={natoms}, found {len(xyz)} atoms")
    else:
        warnings.warn(f"no <sho_atoms> found in grid_Hamiltonian in file \"{filename}\"")
    sp = root.find("spacing")
    for d, a in enumerate("xyz"):
        if sp is not None and sp.get(a):
            hg[d] = float(sp.get(a))
    bnd = root.find("boundary")
    for d, a in enumerate("xyz"):
        if bnd is not None and bnd.get(a):
            bc[d] = int(bnd.get(a))
    pot = root.find("potential")
    if pot is not None:
        for d, a in enumerate(("nx", "ny", "nz")):
            if pot.get(a):
                ng[d] = int(pot.get(a))
        V = _potential(_numbers(pot.text), ng, filename)
    else:
        warnings.warn(f"grid_Hamiltonian has no potential in file \"{filename}\"")
        V = np.zeros((ng[2], ng[1], ng[0]))
    return Hamiltonian(tuple(ng), tuple(hg), tuple(bc), V, np.array(xyz, dtype=np.float64).reshape(-1, 8), hs)


def _load_json(filename):
    with open(filename) as f:
        doc = json.load(f)
    ng, hg, bc = [0, 0, 0], [1.0, 1.0, 1.0], [0, 0, 0]
    if "spacing" in doc:
        hg = [float(v) for v in doc["spacing"]]
    if "boundary" in doc:
        bc = [int(v) for v in doc["boundary"]]
    xyz, hs = [], []
    sa = doc.get("sho_atoms", {})
    for atom in sa.get("atoms", []):
        gid = int(atom.get("atom_id", -1))
        prj = atom.get("projectors", {})
        numax, sigma = int(prj.get("numax", -1)), float(prj.get("sigma", -1.0))
        nc = nsho(numax)
        m = np.zeros((2, nc, nc))
        for k, tag in enumerate(("hamiltonian", "overlap")):
            if tag in atom:
                m[k] = _matrix(np.array(atom[tag], dtype=np.float64), nc, tag, gid, filename)
        p = [float(v) for v in atom.get("position", [0.0, 0.0, 0.0])]
        xyz.append([p[0], p[1], p[2], float(atom.get("Z", 0)), gid, numax, sigma, 0.0])
        hs.append(m)
    if "number" in sa and int(sa["number"]) != len(xyz):
        raise ValueError(f"Inconsistency in file \"{filename}\": number={sa['number']}, found {len(xyz)} atoms")
    pot = doc.get("potential", {})
    if "grid" in pot:
        ng = [int(v) for v in pot["grid"]]
    V = _potential(pot.get("values", []), ng, filename)
    return Hamiltonian(tuple(ng), tuple(hg), tuple(bc), V, np.array(xyz, dtype=np.float64).reshape(-1, 8), hs)


def load_hamiltonian(filename):
    """A grid Hamiltonian from an Hmt.xml or Hmt.json file (a name containing ".json" is read as JSON first,
    falling back to XML, as src/green_input.cxx:43-53 does)."""
    if ".json" in str(filename):
        try:
            return _load_json(filename)
        except (ValueError, KeyError, TypeError):
            pass
    return _load_xml(filename)


def write_hamiltonian(ham, filename, fileformat="xml"):
    """Writes ham in the reference's Hmt format (positions as given, in the centred frame)."""
    nx, ny, nz = ham.ng
    V = np.asarray(ham.V, dtype=np.float64).reshape(-1)
    if fileformat.lower().startswith("x"):
        L = ['<?xml version="1.0"?>', '<grid_Hamiltonian version="0.0">', "  <!-- Units: Hartree and Bohr radii. -->",
             f'  <sho_atoms number="{ham.natoms}">']
        for a in range(ham.natoms):
            x = ham.xyzZinso[a]
            numax = int(x[5])
            L.append(f'    <atom global_id="{int(x[4])}" Z="{int(round(x[3]))}">')
            L.append(f'      <position x="{x[0]:.12f}" y="{x[1]:.12f}" z="{x[2]:.12f}" unit="Bohr"/>')
            L.append(f'      <projectors type="sho" numax="{numax}" sigma="{x[6]:.12f}" unit="Bohr"/>')
            for k, tag in enumerate(("hamiltonian", "overlap")):
                L.append(f'      <{tag} unit="{"1" if k else "Hartree"}">')
                for row in ham.hs[a][k]:
                    L.append("        " + "".join(" 0" if (k and v == 0) else f" {v:.15e}" for v in row))
                L.append(f"      </{tag}>")
            L.append("    </atom>")
        L.append("  </sho_atoms>")
        L.append(f'  <spacing x="{ham.hg[0]:.17f}" y="{ham.hg[1]:.17f}" z="{ham.hg[2]:.17f}" unit="Bohr"/>')
        L.append(f'  <boundary x="{ham.bc[0]}" y="{ham.bc[1]}" z="{ham.bc[2]}"/>')
        head = (f'  <potential nx="{nx}" ny="{ny}" nz="{nz}" minimum="{V.min():g}" maximum="{V.max():g}" '
                f'average="{V.mean():g}" unit="Hartree">')
        body = ["    " + "".join(f" {v:.15f}" for v in V[i:i + 4]) for i in range(0, V.size, 4)]
        L += [head] + body + ["  </potential>", "</grid_Hamiltonian>"]
        text = "\n".join(L) + "\n"
    else:
        atoms = []
        for a in range(ham.natoms):
            x = ham.xyzZinso[a]
            atoms.append({"atom_id": int(x[4]), "Z": int(round(x[3])), "position": [float(v) for v in x[:3]],
                          "projectors": {"type": "sho", "numax": int(x[5]), "sigma": float(x[6])},
                          "hamiltonian": ham.hs[a][0].tolist(), "overlap": ham.hs[a][1].tolist()})
        doc = {"comment": "grid_Hamiltonian in units of Hartree and Bohr radii",
               "sho_atoms": {"number": ham.natoms, "atoms": atoms}, "spacing": list(map(float, ham.hg)),
               "boundary": list(map(int, ham.bc)), "potential": {"grid": [nx, ny, nz], "values": V.tolist()}}
        text = json.dumps(doc, indent=1)
    with open(filename, "w") as f:
        f.write(text)


def from_system(system):
    """The Hamiltonian of a synthetic system (for writing it to a file)."""
    return Hamiltonian(tuple(system.ng), tuple([system.h] * 3), tuple(system.bc), np.asarray(system.V),
                       np.asarray(system.xyzZinso, dtype=np.float64), [np.asarray(m) for m in system.hs])
