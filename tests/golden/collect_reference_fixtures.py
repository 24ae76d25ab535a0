"""Copy the golden data files the reference's own tests hold into
tests/golden/ref/, with a provenance manifest (source path + sha256).

Run where the reference sources are present:
    python tests/golden/collect_reference_fixtures.py [/path/to/reference]

Only data is copied (meshes, gold states); trajectory golds stored as text in
the reference are converted to .npy (float64, the values np.loadtxt parses).
"""
import hashlib
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "ref")

MESH_DIRS = {
    "tests_mesh/mesh_generation_full_1d": ["mesh_s3_5x1_nonperiodic", "mesh_s3_5x1_periodic",
                                           "mesh_s5_5x1_nonperiodic", "mesh_s5_5x1_periodic",
                                           "mesh_s7_5x1_nonperiodic", "mesh_s7_5x1_periodic"],
    "tests_mesh/mesh_generation_full_2d": ["mesh_s3_5x5_periodic_x", "mesh_s3_5x5_periodic_xy",
                                           "mesh_s3_5x5_periodic_y", "mesh_s3_5x6_nonperiodic",
                                           "mesh_s3_5x6_periodic_xy", "mesh_s3_6x5_periodic_xy",
                                           "mesh_s5_7x8_nonperiodic", "mesh_s5_7x8_periodic_x",
                                           "mesh_s5_7x8_periodic_xy", "mesh_s5_8x7_nonperiodic",
                                           "mesh_s5_8x7_periodic_xy", "mesh_s5_8x8_nonperiodic",
                                           "mesh_s5_8x8_periodic_xy", "mesh_s7_7x8_periodic_x",
                                           "mesh_s7_7x8_periodic_xy", "mesh_s7_8x7_periodic_xy",
                                           "mesh_s7_8x8_periodic_xy"],
    "tests_mesh/mesh_generation_full_3d": ["mesh_s3_5x5x5_nonperiodic", "mesh_s3_5x5x5_periodic_xyz",
                                           "mesh_s5_5x5x5_nonperiodic", "mesh_s5_5x5x5_periodic_xyz"],
}
SAMPLE_DIR = "tests_mesh/mesh_generation_sample_2d/mesh_s3_5x6_nonperiodic_sample"

TRAJ = [
    "tests_cpp/eigen_1d_euler_sod_explicit/firstorder/gold.txt",
    "tests_cpp/eigen_1d_euler_sod_explicit/weno3/gold.txt",
    "tests_cpp/eigen_1d_euler_sod_explicit/weno5/gold.txt",
    "tests_cpp/eigen_2d_euler_riemann_explicit/firstorder/rho_gold.txt",
    "tests_cpp/eigen_2d_euler_riemann_explicit/firstorder/p_gold.txt",
    "tests_cpp/eigen_2d_euler_riemann_explicit/weno3/rho_gold.txt",
    "tests_cpp/eigen_2d_euler_riemann_explicit/weno3/p_gold.txt",
    "tests_cpp/eigen_2d_euler_riemann_explicit/weno5/rho_gold.txt",
    "tests_cpp/eigen_2d_euler_riemann_explicit/weno5/p_gold.txt",
    "tests_cpp/eigen_2d_swe_slip_wall_explicit/firstorder/h_gold.txt",
    "tests_cpp/eigen_2d_swe_slip_wall_explicit/weno3/h_gold.txt",
    "tests_cpp/eigen_2d_swe_slip_wall_explicit/weno5/h_gold.txt",
    "tests_cpp/eigen_2d_burgers_periodic_explicit/firstorder/gold.txt",
    "tests_cpp/eigen_2d_burgers_periodic_explicit/weno3/gold.txt",
    "tests_cpp/eigen_2d_burgers_periodic_explicit/weno5/gold.txt",
    "tests_cpp/eigen_3d_euler_sedov_symmetry/firstorder_nxnynz_equal/gold_state.txt",
    "tests_cpp/eigen_3d_euler_sedov_symmetry/weno3_nxnynz_equal/gold_state.txt",
    # implicit (BDF1 with pressio's Newton + BiCGSTAB; Crank-Nicolson)
    "tests_cpp/eigen_1d_euler_sod_implicit/firstorder/gold.txt",
    "tests_cpp/eigen_1d_euler_sod_implicit/weno3/gold.txt",
    "tests_cpp/eigen_2d_euler_riemann_implicit/firstorder/rho_gold.txt",
    "tests_cpp/eigen_2d_euler_riemann_implicit/firstorder/p_gold.txt",
    # advection-diffusion-reaction problem A: RK4 (18x18) and Crank-Nicolson (17x17)
    "tests_cpp/eigen_2d_advdiffreac_probA_explicit/firstorder/gold.txt",
    "tests_cpp/eigen_2d_advdiffreac_probA_explicit/weno3/gold.txt",
    "tests_cpp/eigen_2d_advdiffreac_probA_explicit/weno5/gold.txt",
    "tests_cpp/eigen_2d_advdiffreac_probA_implicit/firstorder/gold.txt",
    "tests_cpp/eigen_2d_advdiffreac_probA_implicit/weno3/gold.txt",
    "tests_cpp/eigen_2d_advdiffreac_probA_implicit/weno5/gold.txt",
    # Burgers outflow: Crank-Nicolson on 20x20 (the reference commits no WENO5 gold)
    "tests_cpp/eigen_2d_burgers_outflow_implicit/firstorder/gold.txt",
    "tests_cpp/eigen_2d_burgers_outflow_implicit/weno3/gold.txt",
]


# the reference's Python tests (tests_py): mesh directory + gold files per
# case; meshes are stored gzip-compressed, golds as float64 .npy
TESTS_PY = [
    ("1d_sod/weno3", "", ["gold.txt"]),
    ("1d_sod/weno5", "", ["gold.txt"]),
    ("1d_shu_osher/weno5", "", ["gold.txt"]),
    ("2d_sedov_full/weno3", "", ["p_gold.txt"]),
    ("2d_sedov_full/weno5", "", ["p_gold.txt"]),
    ("2d_swe_slip_wall/weno3", "", ["h_gold.txt"]),
    ("2d_swe_slip_wall/weno5", "", ["h_gold.txt"]),
    ("2d_burgers_periodic", "mesh", ["gold.txt"]),
    ("2d_advection_diffusion_reaction", "mesh", ["gold1.txt", "gold2.txt"]),
    ("3d_sedov_symmetry/firstorder_nxnynz_equal", "", ["gold_state.txt"]),
    ("3d_sedov_symmetry/weno3_nxnynz_equal", "", ["gold_state.txt"]),
    ("ccu_mesh_s3_5x5", "", []),
]


# tests_mesh/mesh_boundary_cells_tagging_2d: the six meshes its main.cc
# loads and the gold GID lists it holds (written as JSON)
TAGGING = "tests_mesh/mesh_boundary_cells_tagging_2d"
TAGGING_MESHES = ["fullmesh_s3", "fullmesh_s5", "fullmesh_s7", "samplemesh_s3", "samplemesh_s5", "samplemesh_s7"]


def tagging_golds(main_cc):
    """name -> list of GIDs, from the `const std::vector<int> name({...});`
    gold declarations of the test's main.cc"""
    import re
    txt = open(main_cc).read()
    out = {}
    for name, body in re.findall(r"const std::vector<int>\s+(\w+)\s*\(\{([^}]*)\}\);", txt):
        out[name] = [int(v) for v in body.replace("\n", " ").split(",") if v.strip()]
    return out


def sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def main(ref):
    manifest = []
    os.makedirs(OUT, exist_ok=True)
    for base, names in MESH_DIRS.items():
        for n in names:
            src = os.path.join(ref, base, n)
            dst = os.path.join(OUT, "mesh", os.path.basename(base), n)
            os.makedirs(dst, exist_ok=True)
            for f in ("info_gold.dat", "connectivity_gold.dat", "coordinates_gold.dat"):
                shutil.copyfile(os.path.join(src, f), os.path.join(dst, f))
                manifest.append({"file": os.path.relpath(os.path.join(dst, f), HERE),
                                 "source": f"{base}/{n}/{f}", "sha256": sha(os.path.join(src, f))})
    src = os.path.join(ref, SAMPLE_DIR)
    dst = os.path.join(OUT, "mesh", "mesh_generation_sample_2d", os.path.basename(SAMPLE_DIR))
    os.makedirs(dst, exist_ok=True)
    for f in ("info.dat", "connectivity.dat", "coordinates.dat", "sample_mesh_gids.dat",
              "stencil_mesh_gids.dat"):
        shutil.copyfile(os.path.join(src, f), os.path.join(dst, f))
        manifest.append({"file": os.path.relpath(os.path.join(dst, f), HERE),
                         "source": f"{SAMPLE_DIR}/{f}", "sha256": sha(os.path.join(src, f))})
    for rel in TRAJ:
        src = os.path.join(ref, rel)
        name = rel.replace("tests_cpp/", "").replace("/", "__").replace(".txt", ".npy")
        dst = os.path.join(OUT, "traj", name)
        os.makedirs(os.path.dirname(dst), exist_ok=True)
        np.save(dst, np.loadtxt(src))
        manifest.append({"file": os.path.relpath(dst, HERE), "source": rel, "sha256": sha(src),
                         "note": "np.loadtxt of the source text, saved as float64 .npy"})
    import gzip
    for case, meshsub, golds in TESTS_PY:
        dst = os.path.join(OUT, "tests_py", case.replace("/", "__"))
        os.makedirs(dst, exist_ok=True)
        for f in ("info.dat", "connectivity.dat", "coordinates.dat"):
            src = os.path.join(ref, "tests_py", case, meshsub, f)
            out = os.path.join(dst, f + ".gz")
            with open(src, "rb") as fi, gzip.GzipFile(out, "wb", mtime=0) as fo:
                fo.write(fi.read())
            manifest.append({"file": os.path.relpath(out, HERE),
                             "source": f"tests_py/{case}/{meshsub + '/' if meshsub else ''}{f}",
                             "sha256": sha(src), "note": "gzip of the source file"})
        for g in golds:
            src = os.path.join(ref, "tests_py", case, g)
            out = os.path.join(dst, g.replace(".txt", ".npy"))
            np.save(out, np.loadtxt(src))
            manifest.append({"file": os.path.relpath(out, HERE), "source": f"tests_py/{case}/{g}",
                             "sha256": sha(src), "note": "np.loadtxt of the source text, saved as float64 .npy"})
    for n in TAGGING_MESHES:
        dst = os.path.join(OUT, "mesh", "mesh_boundary_cells_tagging_2d", n)
        os.makedirs(dst, exist_ok=True)
        for f in ("info.dat", "connectivity.dat", "coordinates.dat"):
            src = os.path.join(ref, TAGGING, n, f)
            shutil.copyfile(src, os.path.join(dst, f))
            manifest.append({"file": os.path.relpath(os.path.join(dst, f), HERE),
                             "source": f"{TAGGING}/{n}/{f}", "sha256": sha(src)})
    src = os.path.join(ref, TAGGING, "main.cc")
    dst = os.path.join(OUT, "mesh", "mesh_boundary_cells_tagging_2d", "gold_gids.json")
    with open(dst, "w") as f:
        json.dump(tagging_golds(src), f, indent=1)
    manifest.append({"file": os.path.relpath(dst, HERE), "source": f"{TAGGING}/main.cc", "sha256": sha(src),
                     "note": "the gold GID lists declared in the source, as JSON"})
    with open(os.path.join(OUT, "MANIFEST.json"), "w") as f:
        json.dump(manifest, f, indent=1)
    print(f"{len(manifest)} fixtures written to {OUT}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
