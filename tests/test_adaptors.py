"""The C++ adaptors of include/superconga_amd.hpp without a device: every
class instantiates for both of the reference's precisions and links against
the library, and the drop-in `simulation_amd` binary is built and refuses a
missing config the way the reference's binary does."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r'''
#include "superconga_amd.hpp"
namespace sa = superconga_amd;
template class sa::Context<double>;
template class sa::Context<float>;
template class sa::RiccatiSolver<double>;
template class sa::RiccatiSolver<float>;
template class sa::BoundaryConditionSpecular<double>;
template class sa::BoundaryConditionSpecular<float>;
template class sa::ComputeOrderParameter<double>;
template class sa::ComputeCurrent<float>;
template class sa::ComputeFreeEnergy<double>;
template class sa::IntegrationIteratorOzaki<double>;
template class sa::IntegrationIteratorOzaki<float>;
template class sa::Accelerator<double>;
template class sa::ComputeVectorPotential<double>;
template class sa::ComputeVectorPotential<float>;
int main() {
  const double o[3] = {1.0, 2.0, -1.0}, n[3] = {1.5, 2.0, -1.0};
  // |d| = (0.5, 0, 0): L2 numer 0.5, denom sqrt(6), eps 3e-4
  const double r = sa::computeResidual(o, n, 3, sa::ResidualNorm::L2);
  return (r > 0.2041 && r < 0.2042) ? 0 : 1;
}
'''


def test_adaptors_instantiate_and_link(product, tmp_path):
    src = tmp_path / "adaptors.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "adaptors"
    lib = os.path.dirname(product.LIB_PATH)
    subprocess.check_call(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror",
                           "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                           "-L" + lib, "-lsuperconga_amd", "-Wl,-rpath," + lib,
                           "-Wl,-rpath-link,/opt/rocm/lib"])
    # runs without a device: only the host residual is evaluated
    assert subprocess.call([str(exe)]) == 0


def simulation_binary(product):
    path = os.path.join(os.path.dirname(product.LIB_PATH), "simulation_amd")
    if not os.path.exists(path):
        pytest.fail("simulation_amd is not built (python -m superconga_amd.build)")
    return path


def test_simulation_binary_usage(product):
    exe = simulation_binary(product)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode != 0 and "Please provide a config file" in r.stderr
    r = subprocess.run([exe, "/nonexistent/config.json"], capture_output=True, text=True)
    assert r.returncode != 0 and "Could not find/open config file" in r.stderr
