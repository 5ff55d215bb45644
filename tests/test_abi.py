"""CPU checks of the drop-in boundary: the shared library builds, loads and
exports every entry point the header declares (no compute calls)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bellman-cuda.h")
EXT_HEADER = os.path.join(ROOT, "include", "bellman-amd.h")


def header_functions(path=HEADER):
    text = open(path).read()
    return re.findall(r"^bc_error\s+(\w+)\s*\(", text, flags=re.M)


def test_extension_header_matches_binding():
    import bellman_amd as B

    assert sorted(header_functions(EXT_HEADER)) == sorted(B.EXTENSION_SYMBOLS)


def test_header_declares_the_reference_surface():
    import bellman_amd as B

    names = header_functions()
    assert len(names) == 70
    assert sorted(names) == sorted(B.HEADER_SYMBOLS)


def test_library_exports_every_declared_symbol():
    import bellman_amd as B

    if not os.path.exists(B.LIB_PATH):
        pytest.skip("libbellman.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", B.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [n for n in B.HEADER_SYMBOLS + B.EXTENSION_SYMBOLS if n not in exported]
    assert not missing, missing
    lib = B.lib()  # loads without a GPU
    for n in B.HEADER_SYMBOLS:
        assert getattr(lib, n) is not None


def test_struct_layouts_match_c_abi():
    """By-value struct sizes as a C compiler lays out the header's structs."""
    import ctypes

    import bellman_amd as B

    src = r'''
#include <stdio.h>
#include "bellman-cuda.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(ff_grand_product_configuration), sizeof(ff_multiply_by_powers_configuration),
         sizeof(ff_inverse_configuration), sizeof(ff_poly_evaluate_configuration), sizeof(ff_sort_u32_configuration),
         sizeof(generate_permutation_polynomials_configuration), sizeof(msm_configuration), sizeof(ntt_configuration));
  return 0;
}
'''
    import tempfile

    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "s")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        sizes = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    mine = [ctypes.sizeof(t) for t in (B.ff_grand_product_configuration, B.ff_multiply_by_powers_configuration, B.ff_inverse_configuration,
                                        B.ff_poly_evaluate_configuration, B.ff_sort_u32_configuration,
                                        B.generate_permutation_polynomials_configuration, B.msm_configuration, B.ntt_configuration)]
    assert sizes == mine


def test_extension_struct_layouts_match_c_abi():
    import ctypes
    import tempfile

    import bellman_amd as B

    names = ["bc_amd_msm_job", "bc_amd_ntt_job", "bc_amd_pn_job", "bc_amd_job", "bc_amd_job_stats", "bc_amd_msm_tuning", "bc_amd_plan",
             "bc_amd_server"]
    src = "#include <stdio.h>\n#include <stddef.h>\n#include \"bellman-amd.h\"\nint main(void) {\n"
    for n in names:
        src += '  printf("%%zu ", sizeof(%s));\n' % n
    src += '  printf("%zu %zu %zu ", offsetof(bc_amd_job, msm), offsetof(bc_amd_job, ntt), offsetof(bc_amd_job, pn));\n'
    src += '  printf("%zu %zu\\n", offsetof(bc_amd_ntt_job, log_extension_degree), offsetof(bc_amd_job_stats, status));\n  return 0;\n}\n'
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "s")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        got = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    mine = [ctypes.sizeof(getattr(B, n)) for n in names]
    mine += [B.bc_amd_job.msm.offset, B.bc_amd_job.ntt.offset, B.bc_amd_job.pn.offset, B.bc_amd_ntt_job.log_extension_degree.offset,
             B.bc_amd_job_stats.status.offset]
    assert got == mine
