"""Write a copy of the reference's reg-lib/Platform.cpp with a HIP platform into oracle/_ref/.

Test infrastructure (the recipe of `make -C oracle hipref`): the copy adds, at four anchors of the
reference file, the branches a maintainer adds to select libnrhip's platform classes
(niftyreg_amd/plugin/HipPlatform.hpp) -- the constructor's factories, SetGpuIdx's device
selection and CreateOptimiser's device-side conjugate gradient.  The new platform is the
PlatformType value 3 (the reference's enum lists Cpu, Cuda, OpenCl; the header is left as it is,
so every other object of the reference build is used unchanged).  The reference tree is only read.

usage: python patch_platform.py <reference root> <output Platform.cpp>
"""
import sys

HIP = "static_cast<PlatformType>(3)"

INSERTS = [
    # after the CPU kernel factory include: the plugin's classes
    ('#include "CpuKernelFactory.h"\n',
     '#include "CpuKernelFactory.h"\n#include "HipPlatform.hpp"\n#include <type_traits>\n'),
    # constructor: the HIP factories before the unsupported-platform error
    ('    else NR_FATAL_ERROR("Unsupported platform type");\n',
     '    else if (platformType == ' + HIP + ') {\n'
     '        platformName = "HIP";\n'
     '        SetGpuIdx(999);\n'
     '        computeFactory = new HipComputeFactory();\n'
     '        contentCreatorFactory = new HipContentCreatorFactory();\n'
     '        kernelFactory = new HipKernelFactory();\n'
     '        measureCreatorFactory = new HipMeasureCreatorFactory();\n'
     '    }\n'
     '    else NR_FATAL_ERROR("Unsupported platform type");\n'),
    # SetGpuIdx: device selection through the C ABI
    ('    if (platformType == PlatformType::Cpu) {\n        gpuIdx = 999;\n    }\n',
     '    if (platformType == PlatformType::Cpu) {\n        gpuIdx = 999;\n    }\n'
     '    else if (platformType == ' + HIP + ') {\n'
     '        gpuIdx = gpuIdxIn == 999 ? 0 : gpuIdxIn;\n'
     '        NiftyRegHip::Check(nrh_init((int)gpuIdx), "nrh_init");\n'
     '    }\n'),
    # CreateOptimiser: the device-side conjugate gradient over the contents' handles
    ('\n    optimiser->Initialise(controlPointGrid->nvox,',
     '\n    else if (platformType == ' + HIP + ') {\n'
     '        if constexpr (std::is_same<Type, float>::value) {\n'
     '            optimiser = new HipConjugateGradient(dynamic_cast<HipF3dContent&>(con),\n'
     '                                                 conBw ? dynamic_cast<HipF3dContent*>(conBw) : nullptr);\n'
     '        } else {\n'
     '            NR_FATAL_ERROR("The HIP platform computes in single precision");\n'
     '        }\n'
     '        controlPointGridData = static_cast<Type*>(controlPointGrid->data);\n'
     '        transformationGradientData = static_cast<Type*>(con.GetTransformationGradient()->data);\n'
     '        if (conBw) {\n'
     '            controlPointGridDataBw = static_cast<Type*>(controlPointGridBw->data);\n'
     '            transformationGradientDataBw = static_cast<Type*>(conBw->GetTransformationGradient()->data);\n'
     '        }\n'
     '    }\n'
     '    else NR_FATAL_ERROR("Unsupported platform type");\n'
     '\n    optimiser->Initialise(controlPointGrid->nvox,'),
]


def main():
    ref_root, out = sys.argv[1], sys.argv[2]
    text = open(f"{ref_root}/reg-lib/Platform.cpp").read()
    for anchor, replacement in INSERTS:
        if text.count(anchor) != 1:
            sys.exit(f"anchor not found exactly once in Platform.cpp: {anchor!r}")
        text = text.replace(anchor, replacement)
    with open(out, "w") as f:
        f.write(text)


if __name__ == "__main__":
    main()
