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


REF_HEADER = "/root/reference/src/bellman-cuda.h"
STRUCTS = ["bc_stream", "bc_event", "bc_mem_pool", "ff_grand_product_configuration", "ff_multiply_by_powers_configuration",
           "ff_inverse_configuration", "ff_poly_evaluate_configuration", "ff_sort_u32_configuration",
           "generate_permutation_polynomials_configuration", "msm_configuration", "ntt_configuration"]
FIELDS = [("msm_configuration", f) for f in ("bases", "scalars", "results", "log_scalars_count", "h2d_copy_finished", "h2d_copy_finished_callback",
                                             "h2d_copy_finished_callback_data", "d2h_copy_finished", "d2h_copy_finished_callback",
                                             "d2h_copy_finished_callback_data")]
FIELDS += [("ntt_configuration", f) for f in ("inputs", "outputs", "log_values_count", "bit_reversed_inputs", "inverse", "can_overwrite_inputs",
                                              "log_extension_degree", "coset_index", "h2d_copy_finished", "h2d_copy_finished_callback",
                                              "h2d_copy_finished_callback_data", "d2h_copy_finished", "d2h_copy_finished_callback",
                                              "d2h_copy_finished_callback_data")]
FIELDS += [("generate_permutation_polynomials_configuration", f) for f in ("indexes", "scalars", "target", "columns_count", "log_rows_count")]


def _prototypes(path):
    """name -> tuple of parameter types (parameter names and whitespace dropped)."""
    text = re.sub(r"//[^\n]*|/\*.*?\*/", "", open(path).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"^bc_error\s+(\w+)\s*\(([^)]*)\)\s*;", text, flags=re.M):
        params = []
        for p in m.group(2).split(","):
            p = " ".join(p.split())
            if p in ("", "void"):
                continue
            p = re.sub(r"\s*\*\s*", " * ", p).strip()
            toks = p.split()
            if toks[-1] not in ("*",) and len(toks) > 1:
                toks = toks[:-1]  # drop the parameter name
            params.append(" ".join(toks))
        out[m.group(1)] = tuple(params)
    return out


def _layout_program(header):
    src = "#include <stdio.h>\n#include <stddef.h>\n#include \"%s\"\n" % header
    src += "typedef bc_error (*any_fn)(void);\n"
    src += "int main(void) {\n"
    for sname in STRUCTS:
        src += '  printf("%%zu %%zu ", sizeof(%s), _Alignof(%s));\n' % (sname, sname)
    for sname, f in FIELDS:
        src += '  printf("%%zu ", offsetof(%s, %s));\n' % (sname, f)
    src += '  printf("%d %d %d %d\\n", (int)bc_success, (int)bc_error_invalid_value, (int)bc_error_memory_allocation, (int)bc_error_not_ready);\n'
    src += "  return 0;\n}\n"
    return src


@pytest.mark.skipif(not os.path.exists(REF_HEADER), reason="the reference header is only present where the reference tree is")
def test_prototypes_equal_the_reference_header():
    ref, mine = _prototypes(REF_HEADER), _prototypes(HEADER)
    assert len(ref) == 70
    assert ref == mine


@pytest.mark.skipif(not os.path.exists(REF_HEADER), reason="the reference header is only present where the reference tree is")
def test_c_caller_built_against_the_reference_header_links_to_libbellman():
    """A C program compiled against the reference's own header, taking the
    address of every entry point it declares, links against libbellman.so and
    runs (no GPU call); its struct sizes, alignments, field offsets and named
    status values equal those seen through include/bellman-cuda.h."""
    import tempfile

    import bellman_amd as B

    if not os.path.exists(B.LIB_PATH):
        pytest.skip("libbellman.so not built")
    names = sorted(_prototypes(REF_HEADER))
    with tempfile.TemporaryDirectory() as d:
        outs = []
        for inc, hdr in ((os.path.dirname(REF_HEADER), "bellman-cuda.h"), (os.path.join(ROOT, "include"), "bellman-cuda.h")):
            c = os.path.join(d, "layout.c")
            open(c, "w").write(_layout_program(hdr))
            exe = os.path.join(d, "layout")
            subprocess.run(["gcc", "-std=c11", "-I", inc, c, "-o", exe], check=True)
            outs.append(subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split())
        assert outs[0] == outs[1]
        caller = "#include \"bellman-cuda.h\"\n#include <stdio.h>\n"
        caller += "static void *const entry_points[] = {\n" + ",\n".join("  (void *)&%s" % n for n in names) + "\n};\n"
        caller += "int main(void) { printf(\"%zu\\n\", sizeof(entry_points) / sizeof(entry_points[0])); return 0; }\n"
        c = os.path.join(d, "caller.c")
        open(c, "w").write(caller)
        exe = os.path.join(d, "caller")
        libdir = os.path.dirname(B.LIB_PATH)
        subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(REF_HEADER), c, "-o", exe, "-L", libdir, "-lbellman", "-Wl,-rpath," + libdir,
                        "-Wl,-rpath-link,/opt/rocm/lib", "-Wl,--no-as-needed"], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.strip()
        assert out == "70"


def test_no_alloc_limits_match_the_plain_call_footprints():
    """BC_AMD_MSM_NO_ALLOC_MAX_LOG is the largest plain MSM whose temporaries
    (bc_amd_job_footprint with the plain call's window width and chunk) fit
    BC_AMD_MSM_NO_ALLOC_BYTES; host arithmetic only."""
    import ctypes

    import bellman_amd as B
    from oracle import oracle as O

    if not os.path.exists(B.LIB_PATH):
        pytest.skip("libbellman.so not built")
    text = open(EXT_HEADER).read()
    nbytes = eval(re.search(r"#define BC_AMD_MSM_NO_ALLOC_BYTES \((\d+)u << (\d+)\)", text).expand(r"\1 << \2"))
    max_log = int(re.search(r"#define BC_AMD_MSM_NO_ALLOC_MAX_LOG (\d+)", text).group(1))

    def plain_piece(L):  # the plain call's chunk of sorted entries per thread
        return 128 if L >= 24 else 64 if L == 23 else 32 if L >= 21 else 16 if L >= 19 else 8 if L >= 16 else 4

    def footprint(L):
        job = B.bc_amd_job()
        job.kind = B.BC_AMD_JOB_MSM
        job.msm = B.msm_job(1, 1, 1, L, O.msm_window_bits(L), plain_piece(L))
        out = ctypes.c_size_t()
        B.call("bc_amd_job_footprint", ctypes.byref(job), ctypes.byref(out))
        return out.value

    assert footprint(max_log) <= nbytes < footprint(max_log + 1)
