// args.hpp -- the command-line convention of the reference drivers
// (ARGUMENTOS, src/erlbm_common.h:164-200): "--name value" pairs, a value is
// the word after the first occurrence of its name, numbers read with
// atoi / atof, and a default when the name is absent.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "erlbm.h"

namespace drv {

struct Args {
    int argc;
    char **argv;

    const char *find(const char *name) const
    {
        for (int i = 1; i < argc - 1; i++)
            if (std::strcmp(argv[i], name) == 0) return argv[i + 1];
        return nullptr;
    }
    bool has(const char *name) const
    {
        for (int i = 1; i < argc; i++)
            if (std::strcmp(argv[i], name) == 0) return true;
        return false;
    }
    std::string str(const char *name, const std::string &dflt) const
    {
        const char *v = find(name);
        return v ? std::string(v) : dflt;
    }
    double real(const char *name, double dflt) const
    {
        const char *v = find(name);
        return v ? std::atof(v) : dflt;
    }
    int integer(const char *name, int dflt) const
    {
        const char *v = find(name);
        return v ? std::atoi(v) : dflt;
    }
};

// --collision bgk | mrt | reg | erlbm (le_colisao, src/erlbm_common.h:44-54).
inline int collision(const std::string &name)
{
    if (name == "bgk") return ERLBM_COL_BGK;
    if (name == "mrt") return ERLBM_COL_MRT;
    if (name == "reg") return ERLBM_COL_REG;
    if (name == "erlbm") return ERLBM_COL_EXT_REG;
    std::fprintf(stderr, "unknown collision operator: %s (use bgk, mrt, reg or erlbm)\n", name.c_str());
    std::exit(1);
}

inline const char *collision_name(int col)
{
    switch (col) {
    case ERLBM_COL_BGK: return "BGK";
    case ERLBM_COL_MRT: return "MRT";
    case ERLBM_COL_REG: return "RLBM";
    case ERLBM_COL_EXT_REG: return "E-RLBM";
    }
    return "?";
}

// --precision fp64 | fp32 | fp16 | fp32-arith | auto (storage of this build).
inline int precision(const std::string &name)
{
    if (name == "fp64") return ERLBM_FP64;
    if (name == "fp32") return ERLBM_FP32_SHIFTED;
    if (name == "fp16") return ERLBM_FP16_SHIFTED;
    if (name == "fp32-arith") return ERLBM_FP32_SHIFTED_F32;
    if (name == "auto") return ERLBM_FP_AUTO;
    std::fprintf(stderr, "unknown precision: %s (use fp64, fp32, fp16, fp32-arith or auto)\n", name.c_str());
    std::exit(1);
}

inline void check(int rc, const char *what)
{
    if (rc != ERLBM_OK) {
        std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, erlbm_last_error());
        std::exit(2);
    }
}

// C's round() for the step counts, as the drivers compute them.
inline int iround(double x) { return (int)std::llround(x); }

}  // namespace drv
