"""Geometries of the reference's shipped runs (data/*), in bohr."""
ANG_BOHR = 0.529177249  # src/pmodules.f90:9

# data/H2O_CCSD_ccpVDZ: UNITS BOHR, atoms H O H
H2O_GOLDEN = [(1, 0.0, 1.079252144093028, 1.474611055780858), (8, 0.0, 0.0, 0.0),
              (1, 0.0, 1.079252144093028, -1.474611055780858)]
# data/BH_EOM: bohr
BH_GOLDEN = [(5, 2.328142417929234, 0.0, 0.0), (1, 0.0, 0.0, 0.0)]
# data/FH_finiteT_MBPT_*: angstrom
FH_GOLDEN = [(9, 0.9168 / ANG_BOHR, 0.0, 0.0), (1, 0.0, 0.0, 0.0)]
# Config 1 water: O at origin, H at (+-0.7572, 0.5865, 0) angstrom
H2O_CONFIG1 = [(8, 0.0, 0.0, 0.0), (1, 0.7572 / ANG_BOHR, 0.5865 / ANG_BOHR, 0.0),
               (1, -0.7572 / ANG_BOHR, 0.5865 / ANG_BOHR, 0.0)]
