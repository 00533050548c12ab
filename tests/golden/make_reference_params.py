"""Copies the reference's own NACA 0012 case parameter file
(misc/gmsh/naca0012/naca0012_simulation.json, data: the case's parameter
tree in the reference's JSON layout) into tests/golden/ as a fixture for the
driver's parameter-file test.  Run in the container that holds the
reference:  python tests/golden/make_reference_params.py"""
import os
import shutil

SRC = "/root/reference/misc/gmsh/naca0012/naca0012_simulation.json"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "naca0012_simulation.json")

if __name__ == "__main__":
    shutil.copyfile(SRC, DST)
    print("wrote", DST)
