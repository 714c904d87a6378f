"""Builds the thin PyTorch-ROCm extension p2r_amd/_p2r_torch*.so in place:

    python setup_torch_ext.py build_ext --inplace      (run from p2r-tests_amd/, after `make`)

It links lib/libp2r_amd.so (rpath $ORIGIN/../lib) and torch's HIP runtime library; it has no
device code of its own."""
import os

from setuptools import setup
from torch.utils.cpp_extension import BuildExtension, CppExtension

HERE = os.path.dirname(os.path.abspath(__file__))
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

setup(
    name="p2r_amd_torch_ext",
    ext_modules=[
        CppExtension(
            name="p2r_amd._p2r_torch",
            sources=["csrc/p2r_torch.cpp"],
            include_dirs=[os.path.join(ROCM, "include")],
            define_macros=[("__HIP_PLATFORM_AMD__", "1"), ("USE_ROCM", "1")],
            library_dirs=[os.path.join(HERE, "lib")],
            libraries=["p2r_amd", "c10_hip"],
            extra_compile_args=["-O2", "-std=c++17", "-g0"],
            extra_link_args=["-Wl,-rpath,$ORIGIN/../lib", "-s"],
        )
    ],
    cmdclass={"build_ext": BuildExtension},
)
