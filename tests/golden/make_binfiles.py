"""Writes CLEO binary initialisation fixtures with the reference's own Python writers, so that the
readers of cleo_amd/csrc/binfiles.hip are pinned by the layout the reference produces rather than
by a restatement of it.

Runs only where the reference's sources are present (it imports cleopy from /root/reference:
cleopy/writebinary.py and cleopy/gbxboundariesbinary_src/create_gbxboundaries.py, which need only
numpy and struct).  The committed outputs are data: the binary files and an .npz with what was
written into them.

  gridbox boundaries: create_gbxboundaries.write_gridboxboundaries_binary (the reference's
      gridfile writer, fed with cleoconstants.hpp for COORD0), for an uneven 3-D grid given as half
      coordinates, a grid given as [min, max, delta], and a 0-D box;  the expected values are the
      reference's own dimless_gridboxboundaries of the same inputs.
  superdroplets: writebinary.writebinary called with the variables, data types, units and scale
      factors write_initsuperdrops_binary (create_initsuperdrops.py:287-361) passes it, for a 3-D
      population and a 0-D box without coordinates.  (write_initsuperdrops_binary itself imports
      ruamel.yaml, which this image lacks; its attribute generators are the caller's choice and
      are replaced by seeded numpy draws.)

    python tests/golden/make_binfiles.py
"""
import os
import sys

import numpy as np

REFERENCE = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    sys.dont_write_bytecode = True  # nothing may be written under the reference's tree
    sys.path.insert(0, REFERENCE)
    from cleopy import writebinary
    from cleopy.gbxboundariesbinary_src import create_gbxboundaries as cgb

    consts = os.path.join(REFERENCE, "libs", "cleoconstants.hpp")
    COORD0 = cgb.get_COORD0_from_constsfile(consts)
    expected = {"COORD0": np.float64(COORD0)}

    grids = {
        "grid_uneven3d": (np.array([0.0, 20.0, 45.0, 80.0]), np.array([0.0, 25.0, 50.0]),
                          np.array([0.0, 30.0, 70.0, 75.0])),
        "grid_delta3d": ([0.0, 100.0, 25.0], [0.0, 60.0, 20.0], [0.0, 40.0, 20.0]),
        "grid_box0d": (np.array([0.0, 100.0]), np.array([0.0, 100.0]), np.array([0.0, 100.0])),
    }
    for name, (z, x, y) in grids.items():
        cgb.write_gridboxboundaries_binary(os.path.join(HERE, name + ".dat"), z, x, y, consts)
        ndims, idxs, bounds = cgb.dimless_gridboxboundaries(z, x, y, COORD0)
        expected[name + "_ndims"] = np.asarray(ndims, np.uint64)
        expected[name + "_idxs"] = np.asarray(idxs, np.uint32)
        expected[name + "_bounds"] = np.asarray(bounds, np.float64)

    # superdroplets: the argument conventions of create_initsuperdrops.write_initsuperdrops_binary
    R0, RHO0 = 1e-6, 1e5 / (1004.64 * 273.15)
    MASS0 = R0 ** 3 * RHO0
    rng = np.random.default_rng(20261202)
    cases = {"initsupers3d": (12, 3, True), "initsupers0d": (1, 8, False)}
    for name, (ngbx, per, coords) in cases.items():
        n = ngbx * per
        gbx = np.repeat(np.arange(ngbx), per).astype(np.uintc)
        xi = rng.integers(1, 10 ** 12, n).astype(np.uint)
        radius = (rng.lognormal(np.log(10e-6), 0.7, n) / R0).astype(np.double)
        msol = (rng.uniform(1e-19, 1e-16, n) / MASS0).astype(np.double)
        c3, c1, c2 = ((rng.random((3, n)) * 0.05).astype(np.double) if coords
                      else (np.zeros(0), np.zeros(0), np.zeros(0)))
        arrays = [gbx, xi, radius, msol, c3, c1, c2]
        datatypes = [np.uintc, np.uint, np.double, np.double] + [np.double] * 3
        ndata = [len(a) for a in arrays]
        data = []
        for a in arrays:
            data += list(a)
        units = [b" ", b" ", b"m", b"g"] + [b"m"] * 3
        scale_factors = np.asarray([1.0, 1.0, R0, MASS0] + [COORD0] * 3, dtype=np.double)
        metastr = "Variables in this file are Superdroplet attributes: [sdgbxindex, xi, radius, msol"
        metastr += ", coord3, coord1, coord2]" if coords else "]"
        writebinary.writebinary(os.path.join(HERE, name + ".dat"), data, ndata, datatypes, units, scale_factors,
                                metastr)
        for k, a in zip(("sdgbxindex", "xi", "radius", "msol", "coord3", "coord1", "coord2"), arrays):
            expected[f"{name}_{k}"] = np.asarray(a)
    np.savez(os.path.join(HERE, "binfiles_expected.npz"), **expected)


if __name__ == "__main__":
    main()
