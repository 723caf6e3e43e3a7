// wire.h -- a program batch in the rsr_programs_create layout, built on the host program
// by program (the search loop's populations between device steps).
#pragma once
#include <stdint.h>

#include <vector>

#include "../../include/rsr_amd.h"
#include "isa.h"

namespace rsr {

// A program batch in the rsr_programs_create layout, built program by program.
struct Wire {
    std::vector<uint8_t> ops;
    std::vector<uint16_t> args;
    std::vector<int64_t> co{0};
    std::vector<uint8_t> ct;
    std::vector<int64_t> cof{0};
    std::vector<float> cv;
    std::vector<int64_t> vo{0};

    int n() const { return (int)co.size() - 1; }
    // program i of the arrays given (value offsets of the source are per program)
    void add(const uint8_t* s_ops, const uint16_t* s_args, const int64_t* s_co, const uint8_t* s_ct,
             const int64_t* s_cof, const float* s_cv, const int64_t* s_vo, int64_t i) {
        for (int64_t k = s_co[i]; k < s_co[i + 1]; ++k) {
            ops.push_back(s_ops[k]);
            args.push_back(s_args[k]);
        }
        int64_t v = s_vo[i];
        for (int64_t k = s_cof[i]; k < s_cof[i + 1]; ++k) {
            ct.push_back(s_ct[k]);
            const int s = vt_size(s_ct[k]) ? vt_size(s_ct[k]) : 1;
            for (int q = 0; q < s; ++q) cv.push_back(s_cv[v + q]);
            v += s;
        }
        co.push_back((int64_t)ops.size());
        cof.push_back((int64_t)ct.size());
        vo.push_back((int64_t)cv.size());
    }
    void add(const Wire& w, int i) {
        add(w.ops.data(), w.args.data(), w.co.data(), w.ct.data(), w.cof.data(), w.cv.data(), w.vo.data(), i);
    }
    void clear() {
        ops.clear();
        args.clear();
        co.assign(1, 0);
        ct.clear();
        cof.assign(1, 0);
        cv.clear();
        vo.assign(1, 0);
    }
    int64_t len(int i) const { return co[i + 1] - co[i]; }
    int create(rsr_context* ctx, rsr_programs** out) const {
        const uint8_t u8 = 0;
        const uint16_t u16 = 0;
        const float f = 0.0f;
        return rsr_programs_create(ctx, n(), ops.empty() ? &u8 : ops.data(), args.empty() ? &u16 : args.data(),
                                   co.data(), ct.empty() ? &u8 : ct.data(), cof.data(), cv.empty() ? &f : cv.data(),
                                   vo.data(), out);
    }
};


}  // namespace rsr
