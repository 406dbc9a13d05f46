// hlm: drop-in for the reference's hlm_cpu_dp executable (hlm/main_hlm_cpu.cpp,
// hlm/init_hlm.cpp): same flags (plus -w, the disorder width W of V in [-W/2, W/2);
// the reference's stream is W = 2), log_hlm.txt and energy_Chern_No.txt rows, with
// the projected problems of every twist point solved on the GPU.
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <vector>

#include "../../include/hip_diag.h"

using namespace std;

static void usage(const char *target) {
    cout << "Usage: " << target << " -[OPTIONS]" << endl;
    cout << "Options:\n";
    cout << "  -x  lx\n";
    cout << "  -y  ly \n";
    cout << "  -q  q [p/q, p set to 1] \n";
    cout << "  -g  theta mesh griding size (n_mesh)\n";
    cout << "  -b  band to be projected onto\n";
    cout << "  -c  # of disorder configurations (n_sample)\n";
    cout << "  -s  seed to generate seeds\n";
    cout << "  -t  # of threads\n";
    cout << "  -w  disorder width W, V in [-W/2, W/2) (default 2)\n";
    cout << "  -f  first sample (resume a run: samples f..f+c-1, rows appended)\n";
}

int main(int argc, char *argv[]) {
    // defaults of hlm/init_hlm.cpp:20-28
    int lx = 3, ly = 3, n_mesh = 32, n_sample = 1, Q = 3, band = 0, num_threads = 1;
    long first = 0;
    double W = 2.0;
    unsigned long seed = std::chrono::high_resolution_clock::now().time_since_epoch().count();
    int ch, err_flag = 0;
    while ((ch = getopt(argc, argv, "x:y:b:q:g:c:s:h:t:w:f:")) != -1) {
        switch (ch) {
            case 'x': lx = atoi(optarg); break;
            case 'y': ly = atoi(optarg); break;
            case 'b': band = atoi(optarg); break;
            case 'q': Q = atoi(optarg); break;
            case 'g': n_mesh = atoi(optarg); break;
            case 'c': n_sample = atoi(optarg); break;
            case 's': seed = atoi(optarg); break;
            case 't': num_threads = atoi(optarg); break;
            case 'w': W = atof(optarg); break;
            case 'f': first = atol(optarg); break;
            default: err_flag++; break;
        }
    }
    if (err_flag || first < 0) {
        usage(argv[0]);
        exit(2);
    }
    {
        ofstream flog("log_hlm.txt");
        flog << "########################################################" << endl;
        flog << setw(25) << "seed: " << seed << endl;
        flog << setw(25) << "lx: " << lx << endl;
        flog << setw(25) << "ly: " << ly << endl;
        flog << setw(25) << "Q: " << Q << endl;
        flog << setw(25) << "n_mesh: " << n_mesh << endl;
        flog << setw(25) << "n_sample: " << n_sample << endl;
        flog << setw(25) << "n_thread: " << num_threads << endl;
        flog << "########################################################" << endl;
    }
    const long n_phi = (long)lx * ly / Q;
    const double per_sample = 32.0 * (n_mesh + 1) * n_mesh * (double)n_phi * n_phi;
    long spb = (long)(34359738368.0 / per_sample);
    if (spb < 1) spb = 1;
    if (spb > n_sample) spb = n_sample;
    hlm_plan_t plan = nullptr;
    int rc = hlm_plan_create(&plan, lx, ly, Q, band, n_mesh, seed, W, (int)spb, HIP_DIAG_FP64);
    if (rc) {
        cerr << "hlm_plan_create failed with status " << rc << endl;
        return 1;
    }
    ofstream fchern("energy_Chern_No.txt", first > 0 ? ios::app : ios::trunc);
    vector<double> E(spb * n_phi), C(spb * n_phi);
    const double nkp = (double)(n_mesh + 1) * n_mesh;
    for (long s0 = 0; s0 < n_sample; s0 += spb) {
        const int cnt = (int)min(spb, (long)n_sample - s0);
        if ((rc = hlm_plan_load(plan, first + s0, cnt))) break;
        if ((rc = hlm_plan_execute(plan))) break;
        if ((rc = hlm_plan_fetch(plan, E.data(), C.data(), nullptr))) {
            if (rc > 0) rc += (int)s0;   // 1 + the run's sample with a non-finite value
            break;
        }
        if (s0 == 0) {
            double ms[7];
            hlm_plan_stage_times(plan, ms);
            // per theta_1 line (n_mesh solved points), reported per k-point and
            // extrapolated as hlm/main_hlm_cpu.cpp:107-108,171-181 does
            const double diag_line = (ms[0] + ms[1] + ms[2] + ms[3]) / 1e3 / (cnt * nkp) * n_mesh;
            const double chern_line = (ms[4] + ms[5]) / 1e3 / (cnt * nkp) * n_mesh;
            const long total = lround((chern_line * n_mesh + diag_line * (n_mesh + 1)) * n_sample);
            const double tot = total > 0 ? (double)total : (chern_line * n_mesh + diag_line * (n_mesh + 1)) * n_sample;
            cerr << std::left << setw(40) << "diagonalization time/k-point: " << diag_line / (n_mesh + 1) << " s\n";
            cerr << setw(40) << "Chern No. calculation time/k-point: " << chern_line / (n_mesh + 1) << " s\n";
            cerr << setw(40) << "estimated total time:" << std::right << setw(3) << setfill('0') << total / 3600
                 << setw(1) << ":";
            cerr << setw(2) << std::right << setfill('0') << (total % 3600) / 60 << setw(1) << ":";
            cerr << setw(2) << std::right << setfill('0') << total % 60 << endl << setfill(' ');
            cerr << std::left << setw(40) << "estimated # of samples/hour:" << n_sample / (tot / 3600.0) << endl;
            cerr << setw(40) << "# module" << "%(time)" << endl;
            cerr << setw(40) << "diagonalization" << diag_line * (n_mesh + 1) * n_sample / tot * 100.0 << endl;
            cerr << setw(40) << "Chern No. calculation" << chern_line * n_mesh * n_sample / tot * 100.0 << endl;
        }
        // hlm/main_hlm_cpu.cpp:194-196
        for (int s = 0; s < cnt; s++)
            for (long i = 0; i < n_phi; i++)
                fchern << setw(8) << first + s0 + s << " " << setw(15) << setprecision(8) << E[s * n_phi + i] << " "
                       << C[s * n_phi + i] << endl;
    }
    hlm_plan_destroy(plan);
    if (rc) {
        if (rc > 0)
            cerr << "hlm pipeline: sample " << first + rc - 1 << " has a non-finite energy or Chern value" << endl;
        else
            cerr << "hlm pipeline failed with status " << rc << endl;
        return 1;
    }
    return 0;
}
