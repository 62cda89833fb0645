"""The reference's own project headers compile unchanged against this
library's project API (include/sledgehamr/) into the gfx950 kernels.

Each case writes a translation unit into a temporary directory that
#includes the reference's header files where they lie (nothing is copied),
adds what the reference's project class header would add around them
(SLEDGEHAMR_FINISH_SETUP and a class with SLEDGEHAMR_INITIALIZE_PROJECT),
instantiates the project's tagging and probe kernels (the main unit of
shm_project_impl.hpp) and one fused stage kernel with its Rhs (and GW
source), and compiles it with hipcc for gfx950.  Static checks pin the
component ids the macros give (scalars, momenta, then the twelve GW
components).

Covered: projects/MinimalExample/MinimalExample.h (a whole project header),
projects/AxionStrings/kernels_rhs.h and kernels_tagging.h (the string
winding user criterion, te * dt modifiers, a GW source), and
projects/FirstOrderPhaseTransition/kernels_rhs.h and kernels_tagging.h (GW
source with Gradient<2>, te / params[n] modifiers).  Skipped where the
reference tree or hipcc is absent.  CPU only: nothing runs on a GPU.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import tempfile

import pytest

REF = "/root/reference"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

pytestmark = pytest.mark.skipif(
    not os.path.isdir(os.path.join(REF, "projects")) or shutil.which(HIPCC) is None
    and not os.path.exists(HIPCC),
    reason="reference tree or hipcc not available")

FLAGS = ["-std=c++17", "-O3", "-fPIC", "--offload-arch=gfx950",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-ffp-contract=on", "-munsafe-fp-atomics", "-Wno-unused-result",
         "-Wno-unknown-pragmas",
         "-I", os.path.join(ROOT, "include"),
         "-I", os.path.join(ROOT, "include", "sledgehamr"),
         "-I", os.path.join(ROOT, "sledgehamr_amd", "csrc")]


def _ref(path):
    return os.path.join(REF, "projects", path)


CASES = {
    "MinimalExample": (
        f'#include "{_ref("MinimalExample/MinimalExample.h")}"\n'
        "using Cls = MinimalExample::MinimalExample;\n"
        "static_assert(Cls::ShmOps::kNScalars == 4 && Cls::ShmOps::kNFields == 2);\n"
        "static_assert(MinimalExample::Scalar::Pi2 == 3 && MinimalExample::Gw::du_yz == 15);\n"
        "#define SHM_PROJECT_CLASS MinimalExample::MinimalExample\n"
        '#include "shm_project_impl.hpp"\n'
        "template __global__ void shm::stage_kernel<Cls::ShmOps, "
        "shm::PolSsprk<shm::kSsStage3>, shm::Prec::F32C16, 2, false, false, 8, 1, 64>"
        "(const shm::StageArgs);\n",
        []),
    "AxionStrings": (
        "#include <sledgehamr.h>\n"
        f'#include "{_ref("AxionStrings/kernels_rhs.h")}"\n'
        f'#include "{_ref("AxionStrings/kernels_tagging.h")}"\n'
        "namespace AxionStrings {\n"
        "SLEDGEHAMR_FINISH_SETUP\n"
        "class RefCheck : public sledgehamr::Sledgehamr {\n"
        "  public:\n"
        "    SLEDGEHAMR_INITIALIZE_PROJECT(RefCheck)\n"
        "};\n"
        "}  // namespace AxionStrings\n"
        "using Cls = AxionStrings::RefCheck;\n"
        "static_assert(Cls::ShmOps::kNScalars == 4 && Cls::ShmOps::kNGw == 16);\n"
        "#define SHM_PROJECT_CLASS AxionStrings::RefCheck\n"
        '#include "shm_project_impl.hpp"\n'
        "template __global__ void shm::stage_kernel<Cls::ShmOps, "
        "shm::PolSsprk<shm::kSsStage2>, shm::Prec::F32C16, 2, true, false, 8, 1, 64>"
        "(const shm::StageArgs);\n",
        ["-DSHM_GW_MODES=3"]),
    "FirstOrderPhaseTransition": (
        "#include <sledgehamr.h>\n"
        f'#include "{_ref("FirstOrderPhaseTransition/kernels_rhs.h")}"\n'
        f'#include "{_ref("FirstOrderPhaseTransition/kernels_tagging.h")}"\n'
        "namespace FirstOrderPhaseTransition {\n"
        "SLEDGEHAMR_FINISH_SETUP\n"
        "class RefCheck : public sledgehamr::Sledgehamr {\n"
        "  public:\n"
        "    SLEDGEHAMR_INITIALIZE_PROJECT(RefCheck)\n"
        "};\n"
        "}  // namespace FirstOrderPhaseTransition\n"
        "using Cls = FirstOrderPhaseTransition::RefCheck;\n"
        "static_assert(Cls::ShmOps::kNScalars == 2 && Cls::ShmOps::kNGw == 14);\n"
        "static_assert(FirstOrderPhaseTransition::Gw::u_xx == 2);\n"
        "#define SHM_PROJECT_CLASS FirstOrderPhaseTransition::RefCheck\n"
        '#include "shm_project_impl.hpp"\n'
        "template __global__ void shm::stage_kernel<Cls::ShmOps, "
        "shm::PolRkn<1, false>, shm::Prec::F32CS, 2, true, false, 8, 1, 64>"
        "(const shm::StageArgs);\n",
        ["-DSHM_GW_MODES=3"]),
}


def _compile(name, text, extra, tmp):
    src = os.path.join(tmp, f"{name}.hip")
    with open(src, "w") as f:
        f.write(text)
    r = subprocess.run([HIPCC, *FLAGS, *extra, "-c", src, "-o",
                        os.path.join(tmp, f"{name}.o")],
                       capture_output=True, text=True)
    return r.returncode, r.stdout + r.stderr


def test_reference_project_headers_compile_unchanged():
    with tempfile.TemporaryDirectory() as tmp:
        with cf.ThreadPoolExecutor(max_workers=3) as ex:
            futs = {n: ex.submit(_compile, n, t, e, tmp) for n, (t, e) in CASES.items()}
            res = {n: f.result() for n, f in futs.items()}
        for n, (rc, out) in res.items():
            assert rc == 0, f"{n}:\n{out[-4000:]}"
            assert os.path.getsize(os.path.join(tmp, f"{n}.o")) > 0
