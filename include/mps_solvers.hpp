// mps_solvers.hpp -- C++ host surface over the C ABI (mps_amd.h) with the
// shape of the reference classes:
//   MPS_GPU_EXPL  explicit weakly-compressible step (MPS_CAL_expl_v1.0/MPS_GPU_EXPL.h:17-37,
//                 getters of MPS_CAL_/MPS_GPU.h:158-166)
//   MPS_CPU_EXPL  semi-implicit step with the pressure Poisson solve
//                 (MPS_CAL_v4.0/MPS_CPU_EXPL.h:13-33); the name is kept for the drop-in
//                 even though the compute runs on the GPU.
// Initial() reads input.dat and GEO.dat from the working directory, step()
// advances one step, mps_cal() runs to TT and writes out/%04d.out every TOUT.
// Errors are returned as the C ABI's status codes instead of exit().
#pragma once

#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "mps_amd.h"

namespace mps_amd {

class MPS_GPU_EXPL {
public:
    explicit MPS_GPU_EXPL(int precision = 0, int rebin_every_move = 0)
        : precision_(precision), rebin_(rebin_every_move) {}
    virtual ~MPS_GPU_EXPL()
    {
        mps_destroy(ctx_);
        if (log_) fclose(log_);
    }
    MPS_GPU_EXPL(const MPS_GPU_EXPL&) = delete;
    MPS_GPU_EXPL& operator=(const MPS_GPU_EXPL&) = delete;

    // MPS_GPU::Initial (MPS_CAL_expl_v1.0/MPS_GPU.cpp:63-115)
    int Initial(const std::string& dir = ".")
    {
        dir_ = dir;
        if (log_) fclose(log_);
        log_ = fopen((dir + "/out/LOG.txt").c_str(), "wt");   // the run log, as the reference keeps it
        int rc = mps_load_case((dir + "/input.dat").c_str(), &P_);
        if (rc) {
            note("input.dat: could not be read (status %d)\n", rc);
            return rc;
        }
        note("input.dat: dim %d, dp %g m, R0 %g dp, RLAP %g dp, rho %g kg/m3, nu %g m2/s, c %g m/s, "
             "dt_max %g s, TT %g s, TOUT %g s\n",
             P_.dim, P_.dp, P_.re_n_mult, P_.re_lap_mult, P_.rho, P_.nu, P_.cs, P_.dt_max, P_.tt, P_.tout);
        P_.precision = precision_;
        P_.rebin_every_move = rebin_;
        int np = 0, nb2 = 0, nb1 = 0;
        const std::string geo = dir + "/GEO.dat";
        if ((rc = mps_geo_read(geo.c_str(), precision_ == 0, &np, &nb2, &nb1, nullptr, nullptr, nullptr)))
            return rc;
        std::vector<double> pos(3 * (size_t)np), vel(3 * (size_t)np), pr((size_t)np);
        if ((rc = mps_geo_read(geo.c_str(), precision_ == 0, &np, &nb2, &nb1, pos.data(), vel.data(),
                               pr.data())))
            return rc;
        mps_destroy(ctx_);
        ctx_ = nullptr;
        rc = mps_create(&P_, np, nb2, nb1, pos.data(), vel.data(), pr.data(), &ctx_);
        if (rc) {
            note("GEO.dat: %d particles; set-up failed with status %d\n", np, rc);
            return rc;
        }
        double g[4];
        int d[4];
        mps_get_grid(ctx_, g, d);
        note("GEO.dat: %d particles (%d dummy, %d wall, %d fluid); %d x %d x %d cells of %g m\n", np, nb2, nb1,
             np - nb2 - nb1, d[0], d[1], d[2], g[3]);
        return MPS_OK;
    }

    // MPS_GPU_EXPL::step (MPS_CAL_expl_v1.0/MPS_GPU_EXPL.cpp:85-147)
    virtual int step()
    {
        int rc = mps_step(ctx_, 1);
        if (!rc) advance();
        return rc;
    }

    // MPS_GPU_EXPL::mps_cal (MPS_CAL_expl_v1.0/MPS_GPU_EXPL.cpp:24-80)
    int mps_cal(long max_steps = -1)
    {
        if (!ctx_) {
            int rc = Initial(dir_);
            if (rc) return rc;
        }
        int count = 0;
        double threshold = 0.0;
        char name[64];
        const auto t0 = std::chrono::steady_clock::now();
        auto wall = [&t0]() {
            return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        };
        while (d_time_ < P_.tt && (max_steps < 0 || i_step_ < max_steps)) {
            if (d_time_ - threshold >= 0.0) {
                snprintf(name, sizeof name, "/out/%04d.out", count);
                int rc = writeCase(dir_ + name);
                if (rc) return rc;
                if (vtk_) {
                    snprintf(name, sizeof name, "/out/part%04d.vtk", count);
                    if ((rc = writeVtk(dir_ + name))) return rc;
                }
                note("output %04d: step %ld, t = %.6f s, dt = %.3e s, wall %.3f s%s\n", count, i_step_, d_time_,
                     d_dt_, wall(), stepNote().c_str());
                threshold += P_.tout;
                count++;
            }
            int rc = step();
            if (rc) {
                note("step %ld stopped with status %d (t = %.6f s)\n", i_step_, rc, d_time_);
                return rc;
            }
        }
        note("done: %ld steps, t = %.6f s, wall %.3f s\n", i_step_, d_time_, wall());
        return MPS_OK;
    }

    // MPS_GPU::writeCase (MPS_CAL_expl_v1.0/MPS_GPU.cpp:138-152)
    int writeCase(const std::string& path) { return mps_write_out(ctx_, path.c_str()); }

    // MPS_POST's part%04d.vtk for the current state (MPS_POST/MPS_POST.cpp:93-172)
    int writeVtk(const std::string& path) { return mps_write_vtk(ctx_, path.c_str()); }

    // The role of MOTION::doMotion (an empty stub, MPS_CAL_/MOTION.cpp:11-16):
    // flagged particles (by id) move as A sin(2 pi f t) along axis.
    int setMotion(const unsigned char* moving, int axis, double amplitude, double freq)
    {
        return mps_set_motion(ctx_, moving, axis, amplitude, freq);
    }

    // getters (MPS_CAL_/MPS_GPU.h:158-166); arrays refreshed from the device
    int getNp()
    {
        int t = 0, a = 0;
        mps_get_np(ctx_, &t, &a);
        return a;
    }
    const double* getPos() { refresh(); return pos_.data(); }
    const double* getVel() { refresh(); return vel_.data(); }
    const double* getPres() { refresh(); return pr_.data(); }
    const int* getId() { refresh(); return id_.data(); }
    const int* getType() { refresh(); return type_.data(); }

    // also write out/part%04d.vtk next to each .out file in mps_cal()
    void setWriteVtk(bool on) { vtk_ = on; }

    double time() const { return d_time_; }
    long steps() const { return i_step_; }
    mps_ctx* context() { return ctx_; }
    const mps_params& params() const { return P_; }

protected:
    void advance()
    {
        double s[10];
        if (mps_get_scalars(ctx_, s) == MPS_OK) {
            d_time_ = s[4];
            d_dt_ = s[3];
        }
        i_step_++;
        stale_ = true;
    }
    // extra text for the output lines of the log (the semi-implicit step adds its CG figures)
    virtual std::string stepNote() const { return std::string(); }
    // one line to stdout and to out/LOG.txt (when the out folder exists)
    void note(const char* fmt, ...)
    {
        va_list a;
        va_start(a, fmt);
        vprintf(fmt, a);
        va_end(a);
        if (log_) {
            va_start(a, fmt);
            vfprintf(log_, fmt, a);
            va_end(a);
            fflush(log_);
        }
    }
    void refresh()
    {
        if (!stale_ && !pos_.empty()) return;
        int t = 0;
        mps_get_np(ctx_, &t, nullptr);
        pos_.resize(3 * (size_t)t); vel_.resize(3 * (size_t)t); pr_.resize(t); id_.resize(t); type_.resize(t);
        mps_get_state(ctx_, 0, pos_.data(), vel_.data(), pr_.data(), nullptr, id_.data(), type_.data());
        stale_ = false;
    }

    int precision_, rebin_;
    std::string dir_ = ".";
    mps_params P_{};
    mps_ctx* ctx_ = nullptr;
    double d_time_ = 0.0, d_dt_ = 0.0;
    long i_step_ = 0;
    FILE* log_ = nullptr;
    bool stale_ = true;
    bool vtk_ = false;
    std::vector<double> pos_, vel_, pr_;
    std::vector<int> id_, type_;
};

// Semi-implicit step (MPS_CAL_v4.0/MPS_CPU_EXPL.cpp:79-135): PPE with the
// reference's compressibility alpha = 1 / (rho c^2), CG to tol from the
// previous pressure (x0_kind 1) or the learned guess (x0_kind 3).
class MPS_CPU_EXPL : public MPS_GPU_EXPL {
public:
    explicit MPS_CPU_EXPL(int precision = 0, int rebin_every_move = 0, double tol = 1e-8, int x0_kind = 1)
        : MPS_GPU_EXPL(precision, rebin_every_move), tol_(tol), x0_kind_(x0_kind) {}
    int step() override
    {
        const double alpha = 1.0 / (P_.rho * P_.cs * P_.cs);
        int it = 0;
        double ms = 0.0;
        int rc = mps_step_semi(ctx_, 1, alpha, x0_kind_, tol_, &it, &ms);
        if (!rc) {
            advance();
            last_iters_ = it;
            last_solve_ms_ = ms;
        }
        return rc;
    }
    int lastIterations() const { return last_iters_; }
    double lastSolveMs() const { return last_solve_ms_; }

protected:
    std::string stepNote() const override
    {
        char b[96];
        snprintf(b, sizeof b, ", last CG %d iterations in %.3f ms", last_iters_, last_solve_ms_);
        return b;
    }

private:
    double tol_;
    int x0_kind_;
    int last_iters_ = 0;
    double last_solve_ms_ = 0.0;
};

}  // namespace mps_amd
