// hipRTC compile of one specialized pass: the device templates (header) and
// the generated kernel (body) to a gfx950 code object. Shared by the engine
// (spec_pass.hip) and its compile worker (rtc_worker.cpp), so that both build
// the same code from the same text.
#pragma once

#include <hip/hiprtc.h>

#include <string>
#include <vector>

namespace lqcs_rtc {

// True with the code object in `code`, or false with the compiler log in `log`.
inline bool compile_pass(const char* header, const std::string& body, std::vector<char>& code, std::string& log) {
    const std::string src = std::string(header) + "\n" + body;
    hiprtcProgram prog = nullptr;
    if (hiprtcCreateProgram(&prog, src.c_str(), "lqcs_pass.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        log = "hipRTC: cannot create a program";
        return false;
    }
    const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17",
                          // the same code-generation switch as every HIP source of the engine
                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const bool ok = hiprtcCompileProgram(prog, static_cast<int>(sizeof opts / sizeof opts[0]), opts) == HIPRTC_SUCCESS;
    if (ok) {
        size_t n = 0;
        hiprtcGetCodeSize(prog, &n);
        code.resize(n);
        hiprtcGetCode(prog, code.data());
    } else {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        log.assign(n, '\0');
        if (n) hiprtcGetProgramLog(prog, &log[0]);
    }
    hiprtcDestroyProgram(&prog);
    return ok && !code.empty();
}

}  // namespace lqcs_rtc
