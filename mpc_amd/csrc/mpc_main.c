/*
 * mpc_main.c -- the run driver: the replacement for the reference's MPC.c
 * main loop (MPC.c:23-596) on one GPU, written against the library's
 * device-resident API (include/mpc.h) and the legacy file helpers
 * (include/utils.h).
 *
 *   mpc [-r] [-n run] [-d basedir] [-p f64|f32] [-F steps_per_pass] [-R mirror_R]
 *       [-a aperture] [-k dump_every] [-m max_steps]
 *
 * Files (MPC.c:48-112,485-558, byte formats unchanged):
 *   <basedir>/datar.bin, datai.bin      input field, N doubles each
 *   <basedir>%04d/param.bin             parameters (MPC.c:180-203)
 *   <basedir>%04d/{bp1,bp2,temp,spec}.bin appended per dump (doubles)
 *   <basedir>%04d/other.bin             appended "%f\t%f\n" (err, dz)
 *   <basedir>%04d/resdata.bin, resparam.bin checkpoint every 10 dumps, with
 *                                       .bak copies of the append files
 *   <basedir>%04d/out.bin               final field, 2N doubles interleaved
 *
 * Modes: adaptive step doubling + Richardson (the reference's loop,
 * MPC.c:269-383; default) or, with -F, the fixed-step schedule of S steps
 * per pass with the field resident between dumps.  -m stops after that
 * many accepted steps (used to exercise checkpoint/resume).
 */
#define _GNU_SOURCE
#include <getopt.h>
#include <limits.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "../../include/mpc.h"
#include "../../include/utils.h"

static void die(const char *what, int rc) {
    fprintf(stderr, "mpc: %s: %s\n", what, mpc_strerror(rc));
    exit(1);
}

static double *read_doubles(const char *dir, const char *name, long n) {
    FILE *f = opener(dir, name, "rb", 0);
    double *v = (double *)malloc(sizeof(double) * (size_t)n);
    if (!v || fread(v, sizeof(double), (size_t)n, f) != (size_t)n) { fprintf(stderr, "mpc: short read %s\n", name); exit(1); }
    fclose(f);
    return v;
}

/*
 * Checkpoints (MPC.c:521-544, same files and formats).  The field reaches
 * the host asynchronously (mpc_snapshot_begin) while the run goes on; once
 * the copy is complete a writer thread writes resdata.bin, resparam.bin and
 * the .bak copies of the append files, each .bak cut to the size its file
 * had at the checkpoint (the run keeps appending meanwhile).  Everything is
 * written under a temporary name first and renamed at the end, so a run
 * killed mid-write leaves the previous checkpoint in place.
 */
#define NAPPEND 5
static const char *APPEND_NAMES[NAPPEND] = {"bp1.bin", "bp2.bin", "temp.bin", "spec.bin", "other.bin"};

typedef struct {
    char dir[4096];
    const cufftDoubleComplex *A;     /* the pinned snapshot */
    long N;
    double dzz, z, puls_e;
    int signmem, stepn;
    long sizes[NAPPEND];             /* append files' sizes at the checkpoint */
    int rc;
} ckpt_job;

static int write_prefix(const char *from, const char *to, long nbytes) {
    FILE *a = fopen(from, "rb");
    if (!a) return -1;
    FILE *b = fopen(to, "wb");
    if (!b) { fclose(a); return -1; }
    char buf[1 << 16];
    int rc = 0;
    while (nbytes > 0) {
        const size_t want = nbytes < (long)sizeof buf ? (size_t)nbytes : sizeof buf;
        const size_t got = fread(buf, 1, want, a);
        if (got != want || fwrite(buf, 1, got, b) != got) { rc = -1; break; }
        nbytes -= (long)got;
    }
    fclose(a);
    if (fclose(b) != 0) rc = -1;
    return rc;
}

static void *ckpt_writer(void *arg) {
    ckpt_job *j = (ckpt_job *)arg;
    char path[4300], tmp[4400];
    j->rc = 0;
    snprintf(tmp, sizeof tmp, "%s/resdata.bin.tmp", j->dir);
    FILE *f = fopen(tmp, "wb");
    if (!f || fwrite(j->A, sizeof(cufftDoubleComplex), (size_t)j->N, f) != (size_t)j->N) j->rc = -1;
    if (f && fclose(f) != 0) j->rc = -1;
    snprintf(tmp, sizeof tmp, "%s/resparam.bin.tmp", j->dir);
    f = fopen(tmp, "wb");
    if (!f) j->rc = -1;
    else {
        int ok = fwrite(&j->dzz, 8, 1, f) == 1 && fwrite(&j->z, 8, 1, f) == 1 && fwrite(&j->signmem, 4, 1, f) == 1 &&
                 fwrite(&j->puls_e, 8, 1, f) == 1 && fwrite(&j->stepn, 4, 1, f) == 1;
        if (fclose(f) != 0 || !ok) j->rc = -1;
    }
    for (int i = 0; i < NAPPEND && !j->rc; i++) {
        snprintf(path, sizeof path, "%s/%s", j->dir, APPEND_NAMES[i]);
        snprintf(tmp, sizeof tmp, "%s/%s.bak.tmp", j->dir, APPEND_NAMES[i]);
        if (write_prefix(path, tmp, j->sizes[i]) != 0) j->rc = -1;
    }
    if (!j->rc) {
        const char *names[2] = {"resdata.bin", "resparam.bin"};
        for (int i = 0; i < 2 + NAPPEND && !j->rc; i++) {
            if (i < 2) {
                snprintf(path, sizeof path, "%s/%s", j->dir, names[i]);
            } else {
                snprintf(path, sizeof path, "%s/%s.bak", j->dir, APPEND_NAMES[i - 2]);
            }
            snprintf(tmp, sizeof tmp, "%s.tmp", path);
            if (rename(tmp, path) != 0) j->rc = -1;
        }
    }
    if (j->rc) fprintf(stderr, "mpc: writing the checkpoint in %s failed\n", j->dir);
    return NULL;
}

static long file_size(FILE *fp) {
    struct stat st;
    fflush(fp);
    return fstat(fileno(fp), &st) == 0 ? (long)st.st_size : -1;
}

int main(int argc, char **argv) {
    const char *base = "data";
    int resume = 0, nrun = 0, prec = MPC_F64, fixed_S = 0, dump_every = 1, max_steps = -1;
    double mirror_R = 0.0, aperture = 0.0;
    int c;
    while ((c = getopt(argc, argv, "rn:d:p:F:R:a:k:m:")) != -1) {
        switch (c) {
        case 'r': resume = 1; printf("Resuming previous run!\n"); break;
        case 'n': nrun = atoi(optarg); break;
        case 'd': base = optarg; break;
        case 'p': prec = strcmp(optarg, "f32") == 0 ? MPC_F32 : MPC_F64; break;
        case 'F': fixed_S = atoi(optarg); break;
        case 'R': mirror_R = atof(optarg); break;
        case 'a': aperture = atof(optarg); break;
        case 'k': dump_every = atoi(optarg) > 0 ? atoi(optarg) : 1; break;
        case 'm': max_steps = atoi(optarg); break;
        default: fprintf(stderr, "usage: mpc [-r] [-n run] [-d dir] [-p f64|f32] [-F S] [-R R] [-a a] [-k k] [-m n]\n"); return 2;
        }
    }
    char cpath[4096];
    snprintf(cpath, sizeof cpath, "%s%04d", base, nrun);
    char ppath[4200];
    snprintf(ppath, sizeof ppath, "%s/param.bin", cpath);
    mpc_params p;
    int rc = mpc_read_param_bin(ppath, &p);
    if (rc) die("param.bin", rc);
    const long N = (long)p.dims[0] * p.dims[1] * p.dims[2];
    printf("Using directory %s\nMatrix dimensions are: %d x %d x %d.\n", cpath, p.dims[0], p.dims[1], p.dims[2]);
    p.mirror_R = mirror_R;
    p.aperture = aperture;
    p.steps_per_pass = fixed_S > 0 ? fixed_S : 1;
    p.npasses = fixed_S > 0 ? (int)llround(p.dist / p.Cav) : 1;
    if (fixed_S > 0) p.dzz = p.Cav / fixed_S;

    mpc_ctx *ctx = NULL;
    if ((rc = mpc_create(&p, prec, &ctx))) die("mpc_create", rc);

    FILE *fbp1, *fbp2, *ftemp, *fspec, *foth;
    char resd[4300], resp[4300];   /* read on resume; written by the checkpoint writer */
    snprintf(resd, sizeof resd, "%s/resdata.bin", cpath);
    snprintf(resp, sizeof resp, "%s/resparam.bin", cpath);
    cufftDoubleComplex *A = (cufftDoubleComplex *)malloc(sizeof(cufftDoubleComplex) * (size_t)N);
    double dzz = p.dzz, z = 0.0, puls_e = p.puls_e, err = 0.0;
    int signmem = 1, nstep = 0;
    if (resume) {
        /* MPC.c:97-103,144-150 and the resume parameters of :240-252 (read in GPU mode too) */
        fbp1 = opener(cpath, "bp1.bin", "ab", 1);
        fbp2 = opener(cpath, "bp2.bin", "ab", 1);
        ftemp = opener(cpath, "temp.bin", "ab", 1);
        fspec = opener(cpath, "spec.bin", "ab", 1);
        foth = opener(cpath, "other.bin", "a", 1);
        FILE *f = fopen(resd, "rb");
        if (!f || fread(A, sizeof(cufftDoubleComplex), (size_t)N, f) != (size_t)N) { fprintf(stderr, "mpc: cannot read %s\n", resd); return 1; }
        fclose(f);
        f = fopen(resp, "rb");
        int ok = f && fread(&dzz, 8, 1, f) == 1 && fread(&z, 8, 1, f) == 1 && fread(&signmem, 4, 1, f) == 1 &&
                 fread(&puls_e, 8, 1, f) == 1 && fread(&nstep, 4, 1, f) == 1;
        if (f) fclose(f);
        if (!ok) { fprintf(stderr, "mpc: cannot read %s\n", resp); return 1; }
        if ((rc = mpc_upload(ctx, A))) die("upload", rc);
    } else {
        double *Ar = read_doubles(base, "datar.bin", N);
        double *Ai = read_doubles(base, "datai.bin", N);
        for (long i = 0; i < N; i++) { A[i].x = Ar[i]; A[i].y = Ai[i]; }
        free(Ar); free(Ai);
        fbp1 = opener(cpath, "bp1.bin", "wb", 0);
        fbp2 = opener(cpath, "bp2.bin", "wb", 0);
        ftemp = opener(cpath, "temp.bin", "wb", 0);
        fspec = opener(cpath, "spec.bin", "wb", 0);
        foth = opener(cpath, "other.bin", "w", 0);
        if ((rc = mpc_upload(ctx, A))) die("upload", rc);
        if ((rc = mpc_calibrate(ctx))) die("calibrate", rc);   /* MPC.c:206-216 */
    }
    double *temp = (double *)malloc(sizeof(double) * p.dims[2]), *spec = (double *)malloc(sizeof(double) * p.dims[2]);
    double *BP1 = (double *)malloc(sizeof(double) * p.dims[0]), *BP2 = (double *)malloc(sizeof(double) * p.dims[1]);
    printf("Starting main loop on the GPU!\n");
    int done = 0;
    /* checkpoint state: 0 none, 1 snapshot copying, 2 writer thread running */
    int ck_state = 0;
    ckpt_job job;
    pthread_t writer;
    memset(&job, 0, sizeof job);
    snprintf(job.dir, sizeof job.dir, "%s", cpath);
    job.N = N;
    while (z < p.dist && (max_steps < 0 || done < max_steps)) {
        if (fixed_S > 0) {
            /* fixed schedule: dump_every steps resident on the device */
            const int n0 = (int)llround(z / dzz);
            const int total = (int)llround(p.dist / dzz);
            int k = dump_every;
            if (n0 + k > total) k = total - n0;
            if (max_steps >= 0 && done + k > max_steps) k = max_steps - done;
            if (k <= 0) break;
            if ((rc = mpc_run(ctx, n0, k))) die("run", rc);
            z = (n0 + k) * dzz;
            done += k;
            err = 0.0;
        } else {
            int acc = 0;
            double e1 = 0, d1 = 0;
            if ((rc = mpc_run_adaptive(ctx, &z, &dzz, &signmem, 1, &acc, &e1, &d1))) die("run_adaptive", rc);
            if (acc == 0) break;
            err = e1;
            done += 1;
        }
        /* diagnostics of the accepted field (MPC.c:384-512) */
        mpc_diag d;
        memset(&d, 0, sizeof d);
        d.temp = temp; d.spec = spec; d.BP1 = BP1; d.BP2 = BP2;
        if ((rc = mpc_diagnostics(ctx, &d))) die("diagnostics", rc);
        fwrite(BP1, sizeof(double), (size_t)p.dims[0], fbp1); fflush(fbp1);
        fwrite(BP2, sizeof(double), (size_t)p.dims[1], fbp2); fflush(fbp2);
        fwrite(temp, sizeof(double), (size_t)p.dims[2], ftemp); fflush(ftemp);
        fwrite(spec, sizeof(double), (size_t)p.dims[2], fspec); fflush(fspec);
        fprintf(foth, "%f\t%f\n", err, dzz); fflush(foth);
        if (z > p.dist) z = p.dist;
        printf("We are here: %5.2f\n", z * 100.0 / p.dist);
        /* a finished snapshot goes to the writer thread */
        if (ck_state == 1 && mpc_snapshot_poll(ctx) == 1) {
            if ((rc = mpc_snapshot_wait(ctx, &job.A))) die("snapshot", rc);
            if (pthread_create(&writer, NULL, ckpt_writer, &job) != 0) { fprintf(stderr, "mpc: no writer thread\n"); return 1; }
            ck_state = 2;
        }
        if ((nstep % 10) == 0) {
            /* checkpoint (MPC.c:521-544) */
            printf("We are at step %d, writing checkpoint data into files.\n", nstep);
            if (ck_state == 1) {              /* the previous one is still copying: finish it first */
                if ((rc = mpc_snapshot_wait(ctx, &job.A))) die("snapshot", rc);
                ckpt_writer(&job);
                if (job.rc) return 1;
                ck_state = 0;
            }
            if (ck_state == 2) {              /* its writer must be done with the pinned buffer */
                pthread_join(writer, NULL);
                if (job.rc) return 1;
                ck_state = 0;
            }
            if ((rc = mpc_snapshot_begin(ctx))) die("snapshot", rc);
            job.dzz = dzz; job.z = z; job.signmem = signmem; job.puls_e = puls_e; job.stepn = nstep + 1;
            FILE *fa[NAPPEND] = {fbp1, fbp2, ftemp, fspec, foth};
            for (int i = 0; i < NAPPEND; i++) job.sizes[i] = file_size(fa[i]);
            ck_state = 1;
        }
        nstep++;
    }
    if (ck_state == 1) {
        if ((rc = mpc_snapshot_wait(ctx, &job.A))) die("snapshot", rc);
        ckpt_writer(&job);
        if (job.rc) return 1;
    } else if (ck_state == 2) {
        pthread_join(writer, NULL);
        if (job.rc) return 1;
    }
    if ((rc = mpc_download(ctx, A))) die("download", rc);
    FILE *fo = opener(cpath, "out.bin", "wb", 0);
    fwrite(A, sizeof(double), 2 * (size_t)N, fo);   /* MPC.c:558 */
    fclose(fo);
    fclose(fbp1); fclose(fbp2); fclose(ftemp); fclose(fspec); fclose(foth);
    mpc_destroy(ctx);
    mpc_free_params(&p);
    free(A); free(temp); free(spec); free(BP1); free(BP2);
    return 0;
}
