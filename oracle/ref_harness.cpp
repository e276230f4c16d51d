// ref_harness.cpp -- drives the reference's own MPS_CAL_v4.0 CPU solver
// (compiled from its sources in place by oracle/Makefile) to write golden
// vectors.  TEST INFRASTRUCTURE ONLY: it runs in the build container, never
// on the GPU box, and neither the reference sources nor this binary ship.
//
// Usage: ref_v4 <workdir> [--time REPS]
//   <workdir> holds input.dat and GEO.dat in the reference formats
//   (MPS_CAL_v4.0/MPS.cpp:62-300).  The harness runs MPS_CPU::Initial
//   (loadCase, createGeo, calOnce, divideCell), then the per-particle
//   operators and the Poisson build + CG, and writes <workdir>/ref_out.bin.
//   With --time it instead times the reference's explicit operators on that
//   state (oracle/time_vs_ref.py compares them with the oracle's).
//
// The v4.0 tree is compiled with DIM 2 (MPS_CAL_v4.0/common.h:15), so its
// divideCell<2> fills a 9-entry link table.  For 3-D inputs the harness
// replaces that table with the 27-cell stencil before calling any operator;
// the hash, scan and stable scatter of divideCell are dimension-independent.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <sys/stat.h>
#include <unistd.h>

#include "MPS_CPU_EXPL.h"

namespace {

FILE* g_out = nullptr;

void put(const char* name, int kind, long count, const void* data, size_t elem)
{
    char nm[32] = {0};
    std::strncpy(nm, name, 31);
    std::fwrite(nm, 1, 32, g_out);
    std::fwrite(&kind, sizeof(int), 1, g_out);
    std::fwrite(&count, sizeof(long), 1, g_out);
    std::fwrite(data, elem, (size_t)count, g_out);
}
void put_d(const char* n, const std::vector<double>& v) { put(n, 0, (long)v.size(), v.data(), 8); }
void put_i(const char* n, const std::vector<int>& v) { put(n, 1, (long)v.size(), v.data(), 4); }

}  // namespace

class Harness : public MPS_CPU_EXPL {
public:
    int time_reps = 0;

    // Wall time of the reference's own explicit operators on the loaded state:
    // cal_n (its OpenMP loop) and the per-particle d3_LapVel / d3_GradPress,
    // driven here by an OpenMP loop over particles (both only read the state).
    void time_ops()
    {
        const int N = (int)i_np;
        std::vector<double> sink(N);
        auto ms = [](double t0) { return (omp_get_wtime() - t0) * 1e3; };
        double t_n = 0, t_lap = 0, t_grad = 0;
        for (int r = 0; r < time_reps; r++) {
            double t0 = omp_get_wtime();
            cal_n();
            t_n += ms(t0);
            t0 = omp_get_wtime();
#pragma omp parallel for schedule(static)
            for (int i = 0; i < N; i++) {
                const real3 l = d3_LapVel(i);
                sink[i] = l.x + l.y + l.z;
            }
            t_lap += ms(t0);
            t0 = omp_get_wtime();
#pragma omp parallel for schedule(static)
            for (int i = 0; i < N; i++) {
                const real3 g = d3_GradPress(i);
                sink[i] += g.x + g.y + g.z;
            }
            t_grad += ms(t0);
        }
        double chk = 0;
        for (int i = 0; i < N; i++) chk += sink[i];
        std::printf("{\"n\": %d, \"threads\": %d, \"reps\": %d, \"cal_n_ms\": %.4f, \"lapvel_ms\": %.4f, "
                    "\"gradpress_ms\": %.4f, \"checksum\": %.6e}\n",
                    N, omp_get_max_threads(), time_reps, t_n / time_reps, t_lap / time_reps,
                    t_grad / time_reps, chk);
    }

    void run()
    {
        fid_log = std::fopen("./out/LOG.txt", "at");
        Initial();
        const unsigned N = i_np;
        if (i_dim == 3) {
            // 27-cell stencil, ascending cell index (dz, dy, dx = -1..1)
            for (unsigned c = 0; c < i_num_cells; c++) {
                if (i_part_in_cell[c] == 0) continue;
                int k = 1;
                for (int dz = -1; dz <= 1; dz++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++)
                            i_link_cell[c][k++] = c + dz * i_cell_sheet + dy * i_cell_dx + dx;
                i_link_cell[c][0] = 27;
            }
        }
        if (time_reps > 0) { time_ops(); return; }
        std::vector<double> buf(N);
        std::vector<int> ib(N);

        std::vector<double> geo = {(double)i_dim, d_cell_left, d_cell_back, d_cell_bottom,
                                   d_cell_size, (double)i_cell_dx, (double)i_cell_dy,
                                   (double)i_cell_dz, (double)i_num_cells, d_nzero, d_lambda,
                                   d_dt, d_2bydim_over_nzerobylambda};
        put_d("consts", geo);
        for (unsigned i = 0; i < N; i++) ib[i] = (int)i_id[i];
        put_i("id", ib);
        for (unsigned i = 0; i < N; i++) ib[i] = (int)i_type[i];
        put_i("type", ib);
        for (unsigned i = 0; i < N; i++) ib[i] = (int)i_index[i];
        put_i("index_last", ib);
        std::vector<int> cs(i_num_cells), ce(i_num_cells);
        for (unsigned c = 0; c < i_num_cells; c++) { cs[c] = i_cell_start[c]; ce[c] = i_cell_end[c]; }
        put_i("cell_start", cs);
        put_i("cell_end", ce);
        for (unsigned i = 0; i < N; i++) ib[i] = (int)i_cell_list[i];
        put_i("cell_list", ib);
        for (unsigned i = 0; i < N; i++) buf[i] = d3_pos[i].x;
        put_d("x", buf);
        for (unsigned i = 0; i < N; i++) buf[i] = d3_pos[i].y;
        put_d("y", buf);
        for (unsigned i = 0; i < N; i++) buf[i] = d3_pos[i].z;
        put_d("z", buf);
        for (unsigned i = 0; i < N; i++) buf[i] = d_n[i];
        put_d("n_calonce", buf);   // brute-force n of calOnce, carried through the sort

        cal_n();
        for (unsigned i = 0; i < N; i++) buf[i] = d_n[i];
        put_d("n", buf);

        std::vector<double> lx(N), ly(N), lz(N), gx(N), gy(N), gz(N);
        for (unsigned i = 0; i < N; i++) {
            real3 l = d3_LapVel(i);
            lx[i] = l.x; ly[i] = l.y; lz[i] = l.z;
            real3 g = d3_GradPress(i);
            gx[i] = g.x; gy[i] = g.y; gz[i] = g.z;
        }
        put_d("lap_u", lx); put_d("lap_v", ly); put_d("lap_w", lz);
        put_d("grad_x", gx); put_d("grad_y", gy); put_d("grad_z", gz);

        update_dt();
        std::vector<double> dtv = {d_dt};
        put_d("dt_update", dtv);
        d_dt = d_dt_max;  // buildPoisson below uses the calOnce dt (1 / d_one_over_dt)

        buildPoisson();
        std::vector<int> ti, tj;
        std::vector<double> tv;
        for (unsigned i = 0; i < N; i++)
            for (unsigned j = 0; j < N; j++)
                if (A[i][j] != 0.0) { ti.push_back((int)i); tj.push_back((int)j); tv.push_back(A[i][j]); }
        put_i("A_i", ti); put_i("A_j", tj); put_d("A_v", tv);
        for (unsigned i = 0; i < N; i++) buf[i] = b[i];
        put_d("b", buf);
        for (unsigned i = 0; i < N; i++) buf[i] = x[i];
        put_d("x0", buf);
        std::vector<int> bcv(N);
        for (unsigned i = 0; i < N; i++) bcv[i] = bc[i] ? 1 : 0;
        put_i("bc", bcv);

        solvePoisson();
        for (unsigned i = 0; i < N; i++) buf[i] = x[i];
        put_d("x_cg", buf);
        pressCorr();
        for (unsigned i = 0; i < N; i++) buf[i] = d_press[i];
        put_d("p_corr", buf);
    }
};

int main(int argc, char** argv)
{
    if (argc < 2) { std::fprintf(stderr, "usage: ref_v4 <workdir> [--time REPS]\n"); return 2; }
    if (chdir(argv[1]) != 0) { std::perror("chdir"); return 2; }
    mkdir("out", 0755);
    Harness h;
    if (argc >= 4 && !std::strcmp(argv[2], "--time")) {
        // timing: the thread count comes from OMP_NUM_THREADS
        h.time_reps = std::atoi(argv[3]);
        h.run();
        return 0;
    }
    omp_set_num_threads(1);
    g_out = std::fopen("ref_out.bin", "wb");
    if (!g_out) return 2;
    h.run();
    std::fclose(g_out);
    return 0;
}
