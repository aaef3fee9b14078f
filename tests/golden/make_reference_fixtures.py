"""Extract the reference's own netCDF fixtures (Merzbild.jl test/data/*.nc,
HDF5 files) into tests/golden/reference_fixtures.npz with h5dump.

    python tests/golden/make_reference_fixtures.py [/root/reference]

Only data is read: each dataset is dumped as raw little-endian binary and
reshaped to the dataset's dimensions (netCDF order, unlimited time first).
"""
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
H5DUMP = os.environ.get("H5DUMP", "/opt/conda/bin/h5dump")

FILES = {
    # test/test_2species.jl: 400 Ar (3000 K) + 4000 He (360 K), 800 steps, StableRNG(1234)
    "2species": ("2species_seed1234.nc", ["T", "ndens", "np"]),
    # test/test_bkw.jl: 20,000 pseudo-Maxwell Ar particles, 500 steps, moments 4, 6, 8, 10
    "bkw": ("bkw_20k_seed1234.nc", ["moments", "T", "np"]),
    # test/test_1D_couette.jl: 50 cells x 1000 particles, 6000 steps, output every 1000
    "couette": ("couette_0.0005_50_500.0_300.0_1000.nc", ["T", "ndens", "v", "np"]),
    # test/test_1D_couette_varweight.jl: surface properties of the two walls, output every 1000 steps
    "couette_surf": ("couette_0.0005_50_500.0_300.0_1000_vw200to150_surf.nc",
                     ["timestep", "np", "flux_incident", "flux_reflected", "force", "normal_pressure",
                      "shear_pressure", "kinetic_energy_flux"]),
}


def dims(path, name):
    out = subprocess.run([H5DUMP, "-H", "-d", "/" + name, path], capture_output=True, text=True, check=True).stdout
    m = re.search(r"DATASPACE\s+SIMPLE\s+\{\s*\(([^)]*)\)", out)
    return tuple(int(x) for x in m.group(1).split(","))


def dataset(path, name):
    shape = dims(path, name)
    with tempfile.TemporaryDirectory() as td:
        f = os.path.join(td, "d.bin")
        subprocess.run([H5DUMP, "-d", "/" + name, "-b", "LE", "-o", f, path], capture_output=True, check=True)
        a = np.fromfile(f, dtype="<f8")
    return a.reshape(shape)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    out = {}
    for key, (fname, names) in FILES.items():
        path = os.path.join(ref, "test", "data", fname)
        for n in names:
            out[f"{key}_{n}"] = dataset(path, n)
    dst = os.path.join(HERE, "reference_fixtures.npz")
    np.savez_compressed(dst, **out)
    for k, v in out.items():
        print(k, v.shape)


if __name__ == "__main__":
    main()
