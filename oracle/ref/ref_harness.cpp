// ref_harness.cpp -- drives the reference's own Vector/Gauge/CG/BLAS code on
// inputs read from a file and writes every operator output to a file.  Built
// by oracle/ref/Makefile against the reference sources; C++98 (gnu++98) like
// the reference.  Test infrastructure only.
//
// usage: ref_harness ops <in.bin> <out.bin>
//        ref_harness cg  <in.bin> <out.bin>
//        ref_harness time_tm <in.bin> <reps>       (prints seconds per tm())
//        ref_harness time_cg <in.bin> <iters>      (prints seconds per CG iteration)
//        ref_harness proj <in.bin> <out.bin>       (qcd_projectSU33d of the gauge)
//        ref_harness smear <in.bin> <out.bin> <nAPE> <alphaAPE> <nGauss> <alphaGauss>
//        ref_harness prop <in.bin> <out.bin>       (Propagator assembly and rotations)
//        ref_harness time_ape <in.bin> <iters>     (prints seconds per APE step)
//        ref_harness time_gauss <in.bin> <iters>   (prints seconds per Gaussian step)
//
// in.bin (little endian):
//   int32 L[4]; double kappa, mu; int32 twistSign, boundaryT; double tol;
//   int32 maxIter; double gauge[V*72]; double psi[V*24]; double phi[V*24]
// out.bin: repeated records { char name[48]; int64 n; double data[n]; }
#include <qkxTM.h>
#include <lattice_util.h>
#include <blas_qkxTM.h>
#include <solvers.h>
#include <cstdio>
#include <cstring>
#include <vector>

using namespace qkxTM;

static FILE *g_out = 0;

static void put(const char *name, const double *d, long n)
{
    char buf[48];
    memset(buf, 0, 48);
    strncpy(buf, name, 47);
    fwrite(buf, 1, 48, g_out);
    long long nn = n;
    fwrite(&nn, sizeof(nn), 1, g_out);
    fwrite(d, sizeof(double), n, g_out);
}
static void putv(const char *name, Vector &v) { put(name, &(v.M[0][0][0].real()), v.lV4d * 24); }
static void puts1(const char *name, double x) { put(name, &x, 1); }
static void putg(const char *name, Gauge &g) { put(name, &(g.M[0][0][0][0].real()), g.lV4d * 72); }
static void putp(const char *name, Propagator &p) { put(name, &(p.M[0][0][0][0][0].real()), p.lV4d * 288); }

int main(int argc, char **argv)
{
    if (argc < 4) { fprintf(stderr, "usage\n"); return 2; }
    MPI_Init(&argc, &argv);
    const char *mode = argv[1];
    FILE *fi = fopen(argv[2], "rb");
    if (!fi) { fprintf(stderr, "cannot open %s\n", argv[2]); return 2; }
    int L[4], ts, bt, mi;
    double kappa, mu, tol;
    if (fread(L, sizeof(int), 4, fi) != 4) return 3;
    if (fread(&kappa, 8, 1, fi) != 1 || fread(&mu, 8, 1, fi) != 1) return 3;
    if (fread(&ts, 4, 1, fi) != 1 || fread(&bt, 4, 1, fi) != 1) return 3;
    if (fread(&tol, 8, 1, fi) != 1 || fread(&mi, 4, 1, fi) != 1) return 3;

    LatticeInfo info;
    for (int m = 0; m < 4; m++) { info.L[m] = L[m]; info.P[m] = 1; }
    info.Nprocs = 1;
    info.rank = 0;
    // The timing modes run on every rank of the launch, the lattice split in
    // t over P = (1,1,1,N) as the reference drivers do (bin/test_invert.cpp:
    // P[3] = 8, Nprocs and rank from MPI); each rank reads its t-slab of the
    // input.  The other modes run on one rank.
    const bool timing = !strcmp(mode, "time_tm") || !strcmp(mode, "time_cg") || !strcmp(mode, "time_ape") ||
                        !strcmp(mode, "time_gauss");
    int nprocs = 1, myrank = 0;
    MPI_Comm_size(MPI_COMM_WORLD, &nprocs);
    MPI_Comm_rank(MPI_COMM_WORLD, &myrank);
    if (timing && nprocs > 1) {
        if (L[3] % nprocs != 0) { fprintf(stderr, "L[3] not divisible by the number of ranks\n"); return 2; }
        info.P[3] = nprocs;
        info.Nprocs = nprocs;
        info.rank = myrank;
    } else if (nprocs > 1) {
        fprintf(stderr, "mode %s runs on one rank\n", mode);
        return 2;
    }
    info.kappa = kappa;
    info.mu = mu;
    info.twistSign = ts;
    info.boundaryT = bt;
    info.tol = tol;
    info.maxIter = mi;
    int rank = 0;

    Gauge *u = new Gauge(&info);
    Vector *psi = new Vector(&info);
    Vector *phi = new Vector(&info);
    Vector *out = new Vector(&info);
    Vector *tmp = new Vector(&info);
    long V = psi->lV4d;   // this rank's sites: a contiguous t-slab of the lexicographic input
    const long Vg = (long)L[0] * L[1] * L[2] * L[3];
    const long head = 4 * 4 + 2 * 8 + 2 * 4 + 8 + 4;
    const long slab = (long)info.rank * V;
    if (fseek(fi, head + 8 * 72 * slab, SEEK_SET) != 0) return 3;
    if (fread(&(u->M[0][0][0][0].real()), 8, V * 72, fi) != (size_t)(V * 72)) return 3;
    if (fseek(fi, head + 8 * (72 * Vg + 24 * slab), SEEK_SET) != 0) return 3;
    if (fread(&(psi->M[0][0][0].real()), 8, V * 24, fi) != (size_t)(V * 24)) return 3;
    if (fseek(fi, head + 8 * (96 * Vg + 24 * slab), SEEK_SET) != 0) return 3;
    if (fread(&(phi->M[0][0][0].real()), 8, V * 24, fi) != (size_t)(V * 24)) return 3;
    fclose(fi);

    // The drivers apply the boundary condition to the stored links
    // (bin/test_invert_prop.cpp:72, lib/lattice_util.cpp:708-724).
    u->applyBoundaryCondition(&info);

    if (!strcmp(mode, "time_ape") || !strcmp(mode, "time_gauss")) {
        // Smearing cost per step on this lattice (Gauge::APE_smearing,
        // Vector::Gaussian_smearing), alpha as in bin/test_gauss_smearing.cpp:32-34.
        int reps = atoi(argv[3]);
        Gauge *tmpg = new Gauge(&info);
        tmpg->copy(*u);
        double t0 = MPI_Wtime();
        if (!strcmp(mode, "time_ape")) {
            info.NsmearAPE = reps;
            info.alphaAPE = 0.5;
            Gauge *ape = new Gauge(&info);
            ape->APE_smearing(*tmpg, &info);
            delete ape;
        } else {
            info.NsmearGauss = reps;
            info.alphaGauss = 4.0;
            tmp->copy(*psi);
            out->Gaussian_smearing(*tmp, *u, &info);
        }
        double t1 = MPI_Wtime();
        if (info.rank == 0) fprintf(stderr, "%.9e\n", (t1 - t0) / reps);
        delete tmpg;
        MPI_Finalize();
        return 0;
    }

    if (!strcmp(mode, "time_tm") || !strcmp(mode, "time_cg")) {
        int reps = atoi(argv[3]);
        if (!strcmp(mode, "time_tm")) {
            double t0 = MPI_Wtime();
            for (int r = 0; r < reps; r++) out->tm(*psi, *u, &info);
            double t1 = MPI_Wtime();
            printfQKXTM("%.9e\n", (t1 - t0) / reps);
        } else {
            // CG with maxIter = reps: per-iteration cost (its printf goes to stdout).
            info.maxIter = reps;
            info.tol = 1e-300;
            CG cg(&info, *u);
            out->zero();
            double t0 = MPI_Wtime();
            cg(*out, *psi, &info);
            double t1 = MPI_Wtime();
            if (info.rank == 0) fprintf(stderr, "%.9e\n", (t1 - t0) / reps);
        }
        MPI_Finalize();
        return 0;
    }

    g_out = fopen(argv[3], "wb");
    if (!g_out) return 4;

    if (!strcmp(mode, "cg")) {
        // CG::operator() on D^dag D with the given tol; input = D^dag psi
        // (bin/test_invert_prop.cpp:91-94 pattern).
        Vector *b = new Vector(&info);
        b->tmDag(*psi, *u, &info);
        putv("cg_rhs", *b);
        out->zero();
        CG cg(&info, *u);
        cg(*out, *b, &info);
        putv("cg_x", *out);
        tmp->tm(*out, *u, &info);
        putv("cg_Dx", *tmp);
        fclose(g_out);
        MPI_Finalize();
        return 0;
    }

    if (!strcmp(mode, "proj")) {
        // qcd_projectSU33d (lib/blas.cpp:6) on the spatial links of the input
        puts1("plaq_in", u->calculatePlaquette());
        qcd_projectSU33d(*u);
        putg("proj", *u);
        fclose(g_out);
        MPI_Finalize();
        return 0;
    }

    if (!strcmp(mode, "prop")) {
        // Propagator::absorbVector of 12 columns built from psi and phi,
        // column (nu, c2) = (nu + 1) psi + (c2 + 1) i phi, then the rotations
        // to the physical basis (bin/test_seq_source.cpp:95-97, 143-144) and
        // Vector::rescale(2 kappa).
        Propagator *pr = new Propagator(&info);
        Propagator *pm = new Propagator(&info);
        pr->zero();
        for (int nu = 0; nu < 4; nu++)
            for (int c2 = 0; c2 < 3; c2++) {
                for (long iv = 0; iv < V; iv++)
                    for (int mu = 0; mu < 4; mu++)
                        for (int c1 = 0; c1 < 3; c1++)
                            tmp->M[iv][mu][c1] = (double)(nu + 1) * psi->M[iv][mu][c1] +
                                                 Complex(0., (double)(c2 + 1)) * phi->M[iv][mu][c1];
                pr->absorbVector(*tmp, nu, c2);
            }
        putp("prop_in", *pr);
        puts1("prop_norm2", pr->norm2());
        pm->copy(*pr);
        pr->rotateToPhysicalBasePlus();
        putp("prop_plus", *pr);
        pm->rotateToPhysicalBaseMinus();
        putp("prop_minus", *pm);
        out->copy(*psi);
        out->rescale(2 * info.kappa);
        putv("rescale", *out);
        fclose(g_out);
        delete pr;
        delete pm;
        MPI_Finalize();
        return 0;
    }

    if (!strcmp(mode, "smear")) {
        // bin/test_gauss_smearing.cpp:64-85: APE-smear a copy of the gauge, then
        // Gaussian-smear psi with the APE gauge; also psi smeared with the input
        // gauge itself.
        if (argc < 8) { fprintf(stderr, "smear needs nAPE alphaAPE nGauss alphaGauss\n"); return 2; }
        info.NsmearAPE = atoi(argv[4]);
        info.alphaAPE = atof(argv[5]);
        info.NsmearGauss = atoi(argv[6]);
        info.alphaGauss = atof(argv[7]);
        puts1("plaq_in", u->calculatePlaquette());
        Gauge *gtmp = new Gauge(&info);
        Gauge *ape = new Gauge(&info);
        gtmp->copy(*u);
        ape->copy(*u);
        ape->APE_smearing(*gtmp, &info);
        delete gtmp;
        putg("ape", *ape);
        puts1("plaq_ape", ape->calculatePlaquette());
        tmp->copy(*psi);
        out->Gaussian_smearing(*tmp, *ape, &info);
        putv("gauss_ape", *out);
        tmp->copy(*psi);
        out->Gaussian_smearing(*tmp, *u, &info);
        putv("gauss", *out);
        fclose(g_out);
        delete ape;
        MPI_Finalize();
        return 0;
    }

    // ---- every operator on (psi, phi) -------------------------------------
    out->applyDslash(*psi, *u);                putv("applyDslash", *out);
    out->applyDslash_chiral(*psi, *u);         putv("applyDslash_chiral", *out);
    out->applyDslashDag(*psi, *u);             putv("applyDslashDag", *out);
    out->applyDslash_xx(*psi, *u, EO);         putv("applyDslash_xx_EO", *out);
    out->applyDslash_xx(*psi, *u, OE);         putv("applyDslash_xx_OE", *out);
    out->applyDslash_xx_chiral(*psi, *u, EO);  putv("applyDslash_xx_chiral_EO", *out);
    out->applyDslash_xx_chiral(*psi, *u, OE);  putv("applyDslash_xx_chiral_OE", *out);

    out->applyTwistAddDslash(*phi, *psi, &info);          putv("applyTwistAddDslash", *out);
    out->applyTwistAddDslash_chiral(*phi, *psi, &info);   putv("applyTwistAddDslash_chiral", *out);
    out->applyTwistAddDslash_PC(*phi, *psi, &info, OO);   putv("applyTwistAddDslash_PC_OO", *out);
    out->applyTwistAddDslash_PC(*phi, *psi, &info, EE);   putv("applyTwistAddDslash_PC_EE", *out);
    out->applyTwistAddDslash_PC_chiral(*phi, *psi, &info, OO); putv("applyTwistAddDslash_PC_chiral_OO", *out);
    out->applyTwistAddDslash_PC_chiral(*phi, *psi, &info, EE); putv("applyTwistAddDslash_PC_chiral_EE", *out);
    out->applyDeltaAddDslash(*phi, *psi, &info);          putv("applyDeltaAddDslash", *out);
    out->applyTwistDagAddDslashDag(*phi, *psi, &info);    putv("applyTwistDagAddDslashDag", *out);
    out->applyTwist(*psi, &info);              putv("applyTwist", *out);
    out->applyTwistDag(*psi, &info);           putv("applyTwistDag", *out);
    out->applyTwistInv(*psi, &info);           putv("applyTwistInv", *out);
    out->applyTwistInv_chiral(*psi, &info);    putv("applyTwistInv_chiral", *out);
    out->applyTwistDagInv(*psi, &info);        putv("applyTwistDagInv", *out);
    out->applyGamma5(*psi, &info);             putv("applyGamma5", *out);
    out->applyGamma5_chiral(*psi, &info);      putv("applyGamma5_chiral", *out);

    out->tm(*psi, *u, &info);                  putv("tm", *out);
    out->tm_chiral(*psi, *u, &info);           putv("tm_chiral", *out);
    out->tmDag(*psi, *u, &info);               putv("tmDag", *out);
    out->tmDagtm(*psi, *tmp, *u, &info);       putv("tmDagtm", *out);
    out->Wilson(*psi, *u, &info);              putv("Wilson", *out);
    out->tm_xx(*psi, *u, &info, EE);           putv("tm_xx_EE", *out);
    out->tm_xx(*psi, *u, &info, OO);           putv("tm_xx_OO", *out);
    out->tm_xxDagTm_xx(*psi, *u, &info, EE);   putv("tm_xxDagTm_xx_EE", *out);
    out->tm_xxDagTm_xx(*psi, *u, &info, OO);   putv("tm_xxDagTm_xx_OO", *out);
    out->tm_xx_chiral(*psi, *u, &info, EE);    putv("tm_xx_chiral_EE", *out);
    out->tm_xx_chiral(*psi, *u, &info, OO);    putv("tm_xx_chiral_OO", *out);
    out->tm_xxDagTm_xx_chiral(*psi, *u, &info, EE); putv("tm_xxDagTm_xx_chiral_EE", *out);
    out->tm_xxDagTm_xx_chiral(*psi, *u, &info, OO); putv("tm_xxDagTm_xx_chiral_OO", *out);

    puts1("reCdotXdagY", reCdotXdagY(*psi, *phi));
    puts1("norm2", psi->norm2());
    out->copy(*psi); X_eq_X_p_aY(*out, 0.37, *phi);  putv("X_eq_X_p_aY", *out);
    out->copy(*psi); X_eq_aX_p_Y(*out, -1.25, *phi); putv("X_eq_aX_p_Y", *out);

    fclose(g_out);
    delete u; delete psi; delete phi; delete out; delete tmp;
    MPI_Finalize();
    return 0;
}
