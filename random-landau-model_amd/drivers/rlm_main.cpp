// rlm: drop-in for the reference's rlm_cpu_dp / rlm_gpu_* executables
// (rlm/main_rlm_cpu.cpp, rlm/main_rlm_gpu.cu, rlm/init_rlm.cpp): same flags,
// log_rlm.txt, energy_Chern_No.txt rows and stderr timing report, with every
// sample's (g+1)^2 twist mesh assembled, diagonalised and Chern-reduced on the GPU
// through librlm_hip.so.
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
    cout << "  -n  n_phi\n";
    cout << "  -q  quanta concentration \n";
    cout << "  -i  impurity concentration\n";
    cout << "  -g  theta mesh griding size (n_mesh)\n";
    cout << "  -d  disorder potential K-space cut (offhead)\n";
    cout << "  -c  # of disorder configurations(n_sample)\n";
    cout << "  -s  seed to generate seeds\n";
    cout << "  -t  # of threads\n";
    cout << "  -p  precision: 0 = FP64 (default), 1 = mixed (FP32 reduction + FP64 refinement)\n";
    cout << "  -f  first sample (resume a run: samples f..f+c-1, rows appended)\n";
    cout << "  -G  conventions of the single-precision GPU builds (float inputs, energy_Chern_No.dat)\n";
}

int main(int argc, char *argv[]) {
    // defaults of rlm/init_rlm.cpp:20-28
    long n_phi = 32, n_mesh = 32, n_sample = 1, num_threads = 1, off_override = -1, first = 0;
    int precision = HIP_DIAG_FP64, flags = 0;
    bool prec_set = false;
    double q = 1, imp = 16;
    unsigned long seed = std::chrono::high_resolution_clock::now().time_since_epoch().count();
    int ch, err_flag = 0;
    // the reference's optstring is "n:q:i:g:c:s:h:t:"; -d (documented there) is accepted here
    while ((ch = getopt(argc, argv, "n:q:i:g:d:c:s:h:t:p:f:G")) != -1) {
        switch (ch) {
            case 'n': n_phi = atoi(optarg); break;
            case 'q': q = atof(optarg); break;
            case 'i': imp = atof(optarg); break;
            case 'g': n_mesh = atoi(optarg); break;
            case 'd': off_override = atoi(optarg); break;
            case 'c': n_sample = atoi(optarg); break;
            case 's': seed = atoi(optarg); break;
            case 't': num_threads = atoi(optarg); break;
            case 'p':
                precision = atoi(optarg) == 1 ? HIP_DIAG_MIXED : HIP_DIAG_FP64;
                prec_set = true;
                break;
            case 'G': flags |= HIP_DIAG_CONV_GPU_FP32; break;
            case 'f': first = atol(optarg); break;
            default: err_flag++; break;
        }
    }
    if (err_flag || first < 0) {
        usage(argv[0]);
        exit(2);
    }
    (void)num_threads;  // the GPU path has no host worker threads
    const bool gpu_build = (flags & HIP_DIAG_CONV_GPU_FP32) != 0;
    if (gpu_build && !prec_set) precision = HIP_DIAG_MIXED;   // the GPU builds' FP32 solvers
    {
        ofstream flog("log_rlm.txt");
        flog << "########################################################" << endl;
        flog << setw(25) << "seed: " << seed << endl;
        flog << setw(25) << "n_phi: " << n_phi << endl;
        flog << setw(25) << "n_mesh: " << n_mesh << endl;
        flog << setw(25) << "n_sample: " << n_sample << endl;
        flog << setw(25) << "quanta_concentration: " << q << endl;
        flog << setw(25) << "impurity_concentration: " << imp << endl;
        flog << "########################################################" << endl;
    }
    // batch as many samples as fit a 32 GiB budget (eigenvectors + matrices dominate)
    const double per_sample = 32.0 * (n_mesh + 1) * (n_mesh + 1) * (double)n_phi * n_phi;
    long spb = (long)(34359738368.0 / per_sample);
    if (spb < 1) spb = 1;
    if (spb > n_sample) spb = n_sample;

    auto t1 = chrono::high_resolution_clock::now();
    rlm_plan_t plan = nullptr;
    int rc = rlm_plan_create_ex(&plan, n_phi, q, imp, n_mesh, off_override, seed, (int)spb, precision, flags);
    if (rc) {
        cerr << "rlm_plan_create failed with status " << rc << endl;
        return 1;
    }
    auto t2 = chrono::high_resolution_clock::now();
    const double coeff_time = chrono::duration_cast<chrono::microseconds>(t2 - t1).count() / 1.0E6;
    cerr << std::left << setw(40) << "coefficients initialization time: " << coeff_time << " s\n";

    // the GPU builds write energy_Chern_No.dat (rlm/main_rlm_gpu.cu:129), the CPU builds .txt
    ofstream fchern(gpu_build ? "energy_Chern_No.dat" : "energy_Chern_No.txt", first > 0 ? ios::app : ios::trunc);
    vector<double> E(spb * n_phi), C(spb * n_phi);
    const double nkp = (double)(n_mesh + 1) * (n_mesh + 1);
    for (long s0 = 0; s0 < n_sample; s0 += spb) {
        const int cnt = (int)min(spb, n_sample - s0);
        t1 = chrono::high_resolution_clock::now();
        if ((rc = rlm_plan_load(plan, first + s0, cnt))) break;
        t2 = chrono::high_resolution_clock::now();
        const double load_time = chrono::duration_cast<chrono::microseconds>(t2 - t1).count() / 1.0E6;
        if ((rc = rlm_plan_execute(plan))) break;
        if (precision == HIP_DIAG_MIXED) {
            // the mixed path's own diagnostics: near-degenerate runs it could not solve exactly
            long pi[4] = {0, 0, 0, 0}, ri[2] = {0, 0};
            rlm_plan_pair_info(plan, pi);
            rlm_plan_refine_info(plan, ri);
            if (pi[2] || pi[3] || ri[1])
                cerr << "warning: samples " << first + s0 << ".." << first + s0 + cnt - 1 << ": " << pi[2] + pi[3]
                     << " near-degenerate runs unsolved, " << ri[1]
                     << " refinement chunks cut or unfactorised (energies may carry FP32 coupling errors)" << endl;
        }
        if ((rc = rlm_plan_fetch(plan, E.data(), C.data(), nullptr))) {
            if (rc > 0) rc += (int)s0;   // 1 + the run's sample with a non-finite value
            break;
        }
        if (s0 == 0) {
            double ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (rlm_plan_stage_times(plan, ms) == 0) rlm_plan_elapsed_ms(plan, &ms[2]);   // graph mode: total only
            const double pot = (load_time + ms[0] / 1e3) / cnt;
            const double diag = (ms[1] + ms[2] + ms[3] + ms[4] + ms[5]) / 1e3 / (cnt * nkp);
            const double chern = (ms[6] + ms[7]) / 1e3 / (cnt * nkp);
            const double total = coeff_time + (pot + chern * n_mesh * (n_mesh + 1) + diag * nkp) * n_sample;
            cerr << setw(40) << "potential initialization time: " << pot << " s\n";
            cerr << setw(40) << "diagonalization time/k-point: " << diag << " s\n";
            cerr << setw(40) << "Chern No. calculation time/k-point: " << chern << " s\n";
            cerr << setw(40) << "estimated total time:" << std::right << setw(3) << setfill('0')
                 << (int(total)) / 3600 << setw(1) << ":";
            cerr << setw(2) << setfill('0') << ((int(total)) % 3600) / 60 << setw(1) << ":";
            cerr << setw(2) << setfill('0') << (int(total)) % 60 << endl << setfill(' ');
            cerr << std::left << setw(40) << "estimated # of samples/hour:" << n_sample / (total / 3600.0) << endl;
            cerr << setw(40) << "# module" << "%(time)" << endl;
            cerr << setw(40) << "coefficients" << coeff_time / total * 100.0 << "\n";
            cerr << setw(40) << "potential" << pot * n_sample / total * 100.0 << "\n";
            cerr << setw(40) << "diagonalization" << diag * nkp * n_sample / total * 100.0 << endl;
            cerr << setw(40) << "Chern No. calculation" << chern * n_mesh * (n_mesh + 1) * n_sample / total * 100.0
                 << endl;
        }
        // rlm/main_rlm_cpu.cpp:241-242; GPU builds: right-aligned, float values
        // (rlm/main_rlm_gpu.cu:296-297)
        for (int s = 0; s < cnt; s++)
            for (long i = 0; i < n_phi; i++) {
                if (gpu_build)
                    fchern << std::right << setw(8) << first + s0 + s << " " << setw(15) << setprecision(8)
                           << (float)E[s * n_phi + i] << " " << (float)C[s * n_phi + i] << endl;
                else
                    fchern << std::left << setw(8) << first + s0 + s << " " << setw(15) << setprecision(8)
                           << E[s * n_phi + i] << " " << C[s * n_phi + i] << endl;
            }
    }
    rlm_plan_destroy(plan);
    if (rc) {
        if (rc > 0)
            cerr << "rlm pipeline: sample " << first + rc - 1 << " has a non-finite energy or Chern value" << endl;
        else
            cerr << "rlm pipeline failed with status " << rc << endl;
        return 1;
    }
    return 0;
}
