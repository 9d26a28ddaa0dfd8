"""On-disk formats of the reference's System and CollisionDetector.

  export_system / import_system   System::exportSystem / importSystem
                                  (System.cu:1699-1735, 1855-1910): a header
                                  line of element counts, then one CSV line per
                                  rigid body;
  export_matrices                 System::exportMatrices (System.cu:1912-1930):
                                  D, Minv, r, b, k as Matrix Market files;
  export_collisions               CollisionDetector::exportSystem
                                  (CollisionDetector.cu:667-711): one CSV line
                                  per contact;
  output_contact_force_per_collision
                                  System::outputContactForcePerCollision(file)
                                  (System.cu:1672-1697).

Numbers are written as C++ streams write them by default (6 significant
digits, "%g"), so files match the reference's byte for byte at the default
precision; precision=17 writes every double exactly (checkpoint/resume).
These formats carry rigid bodies only: the beam, plate, shell and 2D-body
counts are written as 0 and ignored on import, and export_system and
export_collisions refuse a scene with a shell mesh (export_matrices refuses
it through k, which would need the shell's elastic forces).
"""
from __future__ import annotations

import os

import numpy as np


def _fmt(precision: int):
    f = f"%.{precision}g"
    return lambda x: f % x


# --- pure writers/readers (arrays in, text out) ------------------------------

def write_system(filename, pos, vel, geom, fixed, precision=6):
    """Body lines of System::exportSystem: index, isFixed, p, "1, 0, 0, 0",
    v, then 0 and r for a sphere or 2 and the half extents for a box."""
    g = _fmt(precision)
    n = len(pos)
    with open(filename, "w") as fh:
        fh.write(f"{n}, 0, 0, 0, \n")
        for i in range(n):
            parts = [str(i), str(int(bool(fixed[i])))]
            parts += [g(x) for x in pos[i]]
            parts += ["1", "0", "0", "0"]
            parts += [g(x) for x in vel[i]]
            if geom[i][1] == 0:
                parts += ["0", g(geom[i][0])]
            else:
                parts += ["2"] + [g(x) for x in geom[i]]
            fh.write(", ".join(parts) + ", \n")


def read_system(filename):
    """Rigid bodies of a System::exportSystem file, with the masses
    System::importSystem assigns: 2600 * 4 * 3.14159 * r^3 / 3 for a sphere
    and 1 for a box (System.cu:1895-1899)."""
    with open(filename) as fh:
        head = fh.readline().replace(",", " ").split()
        nb = int(head[0])
        pos, vel, geom, fixed, mass = [], [], [], [], []
        for _ in range(nb):
            tok = fh.readline().replace(",", " ").split()
            fixed.append(int(tok[1]))
            pos.append([float(x) for x in tok[2:5]])
            vel.append([float(x) for x in tok[9:12]])
            shape = int(tok[12])
            if shape == 0:
                r = float(tok[13])
                geom.append([r, 0.0, 0.0])
                mass.append(2600 * 4.0 * 3.14159 * r ** 3.0 / 3.0)
            else:
                geom.append([float(x) for x in tok[13:16]])
                mass.append(1.0)
    return (np.array(pos, np.float64).reshape(-1, 3), np.array(vel, np.float64).reshape(-1, 3),
            np.array(geom, np.float64).reshape(-1, 3), np.array(fixed, np.int32), np.array(mass, np.float64))


def write_matrix_market_coo(filename, I, J, V, rows, cols):
    """A sparse matrix as Matrix Market coordinate real general (1-based),
    values with 17 significant digits."""
    with open(filename, "w") as fh:
        fh.write("%%MatrixMarket matrix coordinate real general\n")
        fh.write(f"{rows} {cols} {len(V)}\n")
        for i, j, v in zip(I, J, V):
            fh.write(f"{int(i) + 1} {int(j) + 1} {float(v):.17g}\n")


def write_matrix_market_array(filename, x):
    """A vector as a Matrix Market dense column."""
    with open(filename, "w") as fh:
        fh.write("%%MatrixMarket matrix array real general\n")
        fh.write(f"{len(x)} 1\n")
        for v in x:
            fh.write(f"{float(v):.17g}\n")


def write_collisions(filename, idA, idB, nap, p, v, geom, precision=6):
    """CollisionDetector::exportSystem lines: index, idA, idB, normal, gap,
    then position, velocity and geometry of A and of B."""
    g = _fmt(precision)
    with open(filename, "w") as fh:
        for i in range(len(idA)):
            a, b = int(idA[i]), int(idB[i])
            parts = [str(i), str(a), str(b)] + [g(x) for x in nap[i]]
            for body in (a, b):
                parts += [g(x) for x in p[body]] + [g(x) for x in v[body]] + [g(x) for x in geom[body]]
            fh.write(", ".join(parts) + ", \n")


def write_contact_forces(filename, idA, idB, loc, normal, friction, nb, precision=6):
    """System::outputContactForcePerCollision(filename) lines."""
    g = _fmt(precision)
    with open(filename, "w") as fh:
        fh.write(f"{len(idA)}, {nb}, 0, 0, 0, \n")
        for i in range(len(idA)):
            parts = [str(i), str(int(idA[i])), str(int(idB[i]))]
            parts += [g(x) for x in loc[i]] + [g(x) for x in normal[i]] + [g(x) for x in friction[i]]
            fh.write(", ".join(parts) + ", \n")


# --- System-level entry points ------------------------------------------------

def _bodies_only(system, what):
    if system.numShellNodes:
        raise NotImplementedError(f"{what} writes rigid bodies only; this scene has a shell mesh")


def export_system(system, filename, precision=6):
    _bodies_only(system, "export_system")
    nb = system.numBodies
    write_system(filename, system.p.reshape(nb, 3), system.v.reshape(nb, 3), system.geometry().reshape(nb, 3),
                 system.fixed(), precision)


def import_system(system, filename):
    """Adds the file's bodies to an uninitialised System (System::importSystem)."""
    pos, vel, geom, fixed, mass = read_system(filename)
    family = -np.ones(len(mass), np.int32)
    system.add_bodies(pos, vel, mass, geom, fixed, family)
    return len(mass)


def export_matrices(system, directory):
    from dvide_amd import system as sysmod

    os.makedirs(directory, exist_ok=True)
    I, J, V, rows, cols = system.jacobian()
    write_matrix_market_coo(os.path.join(directory, "D.mtx"), I, J, V, rows, cols)
    minv = system.inverse_mass()
    n3 = 3 * len(minv)
    idx = np.arange(n3)
    write_matrix_market_coo(os.path.join(directory, "Minv.mtx"), idx, idx, np.repeat(minv, 3), n3, n3)
    write_matrix_market_array(os.path.join(directory, "r.mtx"), system.vector(sysmod.VEC_R))
    write_matrix_market_array(os.path.join(directory, "b.mtx"), system.vector(sysmod.VEC_B))
    write_matrix_market_array(os.path.join(directory, "k.mtx"), system.vector(sysmod.VEC_K))


def export_collisions(system, filename, precision=6):
    _bodies_only(system, "export_collisions")
    nb = system.numBodies
    a, b, nap, _ = system.contacts()
    write_collisions(filename, a, b, nap, system.p.reshape(nb, 3), system.v.reshape(nb, 3),
                     system.geometry().reshape(nb, 3), precision)


def output_contact_force_per_collision(system, filename, precision=6):
    a, b, _, loc = system.contacts()
    normal, friction = system.contactForcePerCollision()
    write_contact_forces(filename, a, b, loc, normal, friction, system.numBodies, precision)
