"""Copies the run inputs of the reference's 2D displacement tests — each
case's specfem_config.yaml, sources file and STATIONS, as they are — into
tests/golden/configs/<case>/ (data only), so that the YAML front-end can be
fed the fixtures' own files unchanged.

Source directories (read as text, in this container only):
  tests/unit-tests/displacement_tests/Newmark/serial/dim2/<case>
for every case whose mesh database is packed by make_specfem2d_traces.py.

Usage: python tests/golden/make_specfem2d_configs.py [/root/reference]
"""

import os
import shutil
import sys

BASE = "tests/unit-tests/displacement_tests/Newmark/serial/dim2"
CASES = ["HomogeneousElasticDomain", "HomogeneousElasticDomainStaceyDirichletBC",
         "HomogeneousElasticDomainStaceyBC", "HomogeneousElasticDomainWAdjacencyMap",
         "HomogeneousElasticDomainMomentTensor", "HomogeneousElasticDomainMomentTensorHeaviside",
         "HomogeneousAcousticDomain", "HomogeneousAcousticDomainDirichletBC",
         "HomogeneousAcousticDomainCompositeStaceyDirichletBC"]
# AcousticElasticCoupledDomainVariant2 is left out: the reference's test list
# comments it out (dim2/tests.yaml) and its config names the sources under
# `databases`, where the reference's parser does not read them.
FILES = ("specfem_config.yaml", "STATIONS")
# the sources file is named by each config (sources.yaml or source.yaml)
SOURCE_FILES = ("sources.yaml", "source.yaml")


def main(ref_root):
    here = os.path.dirname(os.path.abspath(__file__))
    for case in CASES:
        out = os.path.join(here, "configs", case)
        os.makedirs(out, exist_ok=True)
        src = os.path.join(ref_root, BASE, case)
        names = list(FILES) + [f for f in SOURCE_FILES if os.path.exists(os.path.join(src, f))]
        for f in names:
            shutil.copyfile(os.path.join(src, f), os.path.join(out, f))
        print(case, names)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
