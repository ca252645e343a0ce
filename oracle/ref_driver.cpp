/*
 * ref_driver.cpp -- drives QuadRay's own CPU renderer, compiled from the
 * sources under /root/reference by oracle/Makefile into oracle/_ref/.
 * TEST INFRASTRUCTURE: produces reference frames for the golden fixtures and
 * the multithreaded CPU baseline timed by bench.py.
 *
 *   ref_driver <config> <seed> [-x W -y H] [-a fsaa] [-d depth] [-i]
 *              [-t threads] [-f frames] [-o out.bin] [-q spp]
 *
 * -i renders first-hit ids (every surface side gets a LIGHT material whose
 *    colour encodes its id), -q runs the path tracer with spp samples,
 * -s n,t,k selects QuadRay's SIMD target simd_init(n, t, k) (native size in
 *    128-bit quads, sub-type, size factor; core/engine/engine.cpp:602-657)
 *    instead of the one it picks for the host,
 * -c "T:a,a;T:a" camera actions: after a first render at time 0, for each
 *    group apply rt_Scene::update(T, a) for each action a, then render(T);
 *    the frame written is the last one (frames/warm-up ignored)
 * -r times the render alone: set_opts(RT_OPTS_FULL | RT_OPTS_UPDATE_EXT0)
 *    runs the update stage on the first frame only (core/engine/format.h:67,
 *    core/engine/engine.cpp:3018-3020); without it every frame updates.
 * Configs 101..118 are QuadRay's own test scenes test/scenes/scn_test01..18
 * (the seed is ignored; 800x480, no AA and depth 10 unless given).
 * Prints one JSON line with timings; writes the ARGB8 frame to -o.
 */
#include <pthread.h>
#include <sched.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <vector>

/* the render depth is a private member of rt_Scene; open it for this driver */
#define private public
#include "engine.h"
#undef private

#include "scenegen.hpp"

/* QuadRay's own test scenes, as its core_test builds them
 * (test/core_test.cpp:193-480) */
#include "scn_test01.h"
#include "scn_test02.h"
#include "scn_test03.h"
#include "scn_test04.h"
#include "scn_test05.h"
#include "scn_test06.h"
#include "scn_test07.h"
#include "scn_test08.h"
#include "scn_test09.h"
#include "scn_test10.h"
#include "scn_test11.h"
#include "scn_test12.h"
#include "scn_test13.h"
#include "scn_test14.h"
#include "scn_test15.h"
#include "scn_test16.h"
#include "scn_test17.h"
#include "scn_test18.h"

static rt_SCENE *const test_scenes[18] = {
    &scn_test01::sc_root, &scn_test02::sc_root, &scn_test03::sc_root,
    &scn_test04::sc_root, &scn_test05::sc_root, &scn_test06::sc_root,
    &scn_test07::sc_root, &scn_test08::sc_root, &scn_test09::sc_root,
    &scn_test10::sc_root, &scn_test11::sc_root, &scn_test12::sc_root,
    &scn_test13::sc_root, &scn_test14::sc_root, &scn_test15::sc_root,
    &scn_test16::sc_root, &scn_test17::sc_root, &scn_test18::sc_root,
};

static rt_pntr sys_alloc(rt_size size) { return malloc(size); }
static rt_void sys_free(rt_pntr ptr, rt_size size) { free(ptr); }

/* barrier thread pool: one worker per requested thread, each runs one slice
 * of update_slice / render_slice per command */
struct Pool
{
    rt_Platform *pfm = nullptr;
    int thnum = 1;
    volatile int cmd = 0;
    volatile int quit = 0;
    pthread_barrier_t b0, b1;
    std::vector<pthread_t> thr;
    std::vector<int> idx;
};

static void *worker(void *p)
{
    Pool *pool = ((Pool **)p)[0];
    int ti = (int)(intptr_t)((void **)p)[1];
    free(p);
    for (;;)
    {
        pthread_barrier_wait(&pool->b0);
        if (pool->quit) break;
        rt_Scene *scn = pool->pfm->get_cur_scene();
        int cmd = pool->cmd;
        if ((cmd & 3) == 1) scn->update_slice(ti, cmd >> 2);
        else if ((cmd & 3) == 2) scn->render_slice(ti, cmd >> 2);
        pthread_barrier_wait(&pool->b1);
    }
    pthread_barrier_wait(&pool->b1);
    return nullptr;
}

static Pool *g_pool = nullptr;
static int g_threads = 1;

static rt_pntr init_threads(rt_si32 thnum, rt_Platform *pfm)
{
    Pool *pool = new Pool();
    thnum = thnum < 0 ? -thnum : thnum;
    pool->pfm = pfm;
    pool->thnum = thnum;
    pthread_barrier_init(&pool->b0, nullptr, thnum + 1);
    pthread_barrier_init(&pool->b1, nullptr, thnum + 1);
    pool->thr.resize(thnum);
    for (int i = 0; i < thnum; i++)
    {
        void **arg = (void **)malloc(2 * sizeof(void *));
        arg[0] = pool;
        arg[1] = (void *)(intptr_t)i;
        pthread_create(&pool->thr[i], nullptr, worker, arg);
    }
    pfm->set_thnum(thnum);
    g_pool = pool;
    return pool;
}

static rt_void term_threads(rt_pntr tdata, rt_si32 thnum)
{
    Pool *pool = (Pool *)tdata;
    pool->quit = 1;
    pthread_barrier_wait(&pool->b0);
    pthread_barrier_wait(&pool->b1);
    for (pthread_t t : pool->thr) pthread_join(t, nullptr);
    delete pool;
}

static rt_void run_cmd(rt_pntr tdata, int cmd)
{
    Pool *pool = (Pool *)tdata;
    pool->cmd = cmd;
    pthread_barrier_wait(&pool->b0);
    pthread_barrier_wait(&pool->b1);
}

static rt_void update_scene(rt_pntr tdata, rt_si32 thnum, rt_si32 phase)
{
    run_cmd(tdata, (phase << 2) | 1);
}

static rt_void render_scene(rt_pntr tdata, rt_si32 thnum, rt_si32 phase)
{
    run_cmd(tdata, (phase << 2) | 2);
}

static double now_s()
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

int main(int argc, char **argv)
{
    if (argc < 3)
    {
        fprintf(stderr, "usage: ref_driver config seed [options]\n");
        return 2;
    }
    int config = atoi(argv[1]);
    unsigned long long seed = strtoull(argv[2], nullptr, 10);
    int x_res = -1, y_res = -1, fsaa = -1, depth = -1, ids = 0, fast = 0, frames = 1;
    int spp = -1, warm = 0, simd_req = 0, render_only = 0;
    const char *out = nullptr;
    const char *planes = nullptr;
    const char *actions = nullptr;
    for (int i = 3; i < argc; i++)
    {
        if (!strcmp(argv[i], "-x") && i + 1 < argc) x_res = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-y") && i + 1 < argc) y_res = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-a") && i + 1 < argc) fsaa = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-d") && i + 1 < argc) depth = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-t") && i + 1 < argc) g_threads = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-f") && i + 1 < argc) frames = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-w") && i + 1 < argc) warm = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-q") && i + 1 < argc) spp = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-o") && i + 1 < argc) out = argv[++i];
        else if (!strcmp(argv[i], "-P") && i + 1 < argc) planes = argv[++i];
        else if (!strcmp(argv[i], "-s") && i + 1 < argc)
        {
            int n = 0, t = 0, k = 0;
            if (sscanf(argv[++i], "%d,%d,%d", &n, &t, &k) != 3)
            {
                fprintf(stderr, "ref_driver: -s takes n,t,k\n");
                return 2;
            }
            simd_req = simd_init(n, t, k);
        }
        else if (!strcmp(argv[i], "-i")) ids = 1;
        else if (!strcmp(argv[i], "-p")) fast = 1;
        else if (!strcmp(argv[i], "-r")) render_only = 1;
        else if (!strcmp(argv[i], "-c") && i + 1 < argc) actions = argv[++i];
    }

    qrgen::Scene *sc = nullptr;
    rt_SCENE *root = nullptr;
    if (config >= 101 && config <= 118)
    {
        /* a test scene keeps its own option bits; parity mode adds the
         * pruning options to those it turns off */
        if (ids)
        {
            fprintf(stderr, "ref_driver: -i is not available for test scenes\n");
            return 2;
        }
        root = test_scenes[config - 101];
        if (!fast) root->opts |= RT_OPTS_PRUNING;
        if (x_res < 0) x_res = 800;
        if (y_res < 0) y_res = 480;
        if (fsaa < 0) fsaa = 0;
        if (depth < 0) depth = RT_STACK_DEPTH;
    }
    else
    {
        sc = qrgen::build(config, seed);
        if (ids) qrgen::override_ids(*sc);
        /* fast mode: give QuadRay back its list pruning, tiling and
         * shadow/2-sided culling (its default options); frames then differ
         * from the parity mode by the reference's own tolerance, so this
         * mode is for timing only */
        if (fast) sc->root.opts &= ~RT_OPTS_PRUNING;
        root = &sc->root;
        if (x_res < 0) x_res = sc->x_res;
        if (y_res < 0) y_res = sc->y_res;
        if (fsaa < 0) fsaa = ids ? 0 : sc->fsaa;
        if (depth < 0) depth = sc->depth;
    }
    if (spp < 0) spp = 0;

    rt_Platform *pfm;
    if (g_threads > 1)
    {
        pfm = new rt_Platform(sys_alloc, sys_free, g_threads, init_threads,
                              term_threads, update_scene, render_scene);
    }
    else
    {
        pfm = new rt_Platform(sys_alloc, sys_free);
    }
    pfm->set_fsaa(fsaa);
    int simd = pfm->set_simd(simd_req);

    int x_row = (x_res + 63) / 64 * 64;
    std::vector<rt_ui32> frame_buf((size_t)x_row * y_res + 64);
    rt_Scene *scn = new (pfm) rt_Scene(root, x_res, y_res, x_row, nullptr, pfm);
    scn->depth = (rt_ui32)depth;
    if (render_only) scn->set_opts(RT_OPTS_FULL | RT_OPTS_UPDATE_EXT0);
    if (spp > 0) scn->set_pton(1);

    std::vector<double> times;
    if (actions != nullptr)
    {
        /* camera actions between renders, as RooT drives its camera
         * (root/RooT.h:539-547) */
        scn->render(0);
        const char *p = actions;
        while (*p != '\0')
        {
            char *end = nullptr;
            long long t = strtoll(p, &end, 10);
            if (end == p || *end != ':') { fprintf(stderr, "bad -c\n"); return 2; }
            p = end + 1;
            for (;;)
            {
                long a = strtol(p, &end, 10);
                if (end == p) { fprintf(stderr, "bad -c\n"); return 2; }
                scn->update((rt_time)t, (rt_si32)a);
                p = end;
                if (*p == ',') { p++; continue; }
                break;
            }
            scn->render((rt_time)t);
            if (*p == ';') p++;
            else if (*p != '\0') { fprintf(stderr, "bad -c\n"); return 2; }
        }
        frames = 0;
        warm = 0;
    }
    int total = warm + frames * (spp > 0 ? spp : 1);
    for (int f = 0; f < total; f++)
    {
        double t0 = now_s();
        scn->render(f);
        double t1 = now_s();
        if (f >= warm) times.push_back(t1 - t0);
    }
    std::sort(times.begin(), times.end());
    double med = times.empty() ? 0.0 : times[times.size() / 2];
    double sum = 0.0;
    for (double t : times) sum += t;

    if (getenv("QR_REF_DUMP_CAM") != nullptr)
    {
        rt_Camera *c = scn->cam;
        fprintf(stderr, "pos %.9g %.9g %.9g\n", c->pos[0], c->pos[1], c->pos[2]);
        fprintf(stderr, "hor %.9g %.9g %.9g\n", c->hor[0], c->hor[1], c->hor[2]);
        fprintf(stderr, "ver %.9g %.9g %.9g\n", c->ver[0], c->ver[1], c->ver[2]);
        fprintf(stderr, "nrm %.9g %.9g %.9g\n", c->nrm[0], c->nrm[1], c->nrm[2]);
        fprintf(stderr, "pov %.9g aspect %.9g factor %.9g\n", c->pov,
                scn->aspect, scn->factor);
    }

    /* path tracer: colour planes as [y][x][sample][rgb] floats; QuadRay keeps
     * slot (y * x_row + x) * samples + k of each plane */
    if (planes != nullptr && scn->ptr_r != nullptr)
    {
        int ns_ = fsaa == 2 ? 4 : fsaa == 1 ? 2 : 1;
        FILE *fp = fopen(planes, "wb");
        if (fp == nullptr) { perror(planes); return 1; }
        for (int y = 0; y < y_res; y++)
        {
            for (int x = 0; x < x_res; x++)
            {
                for (int k = 0; k < ns_; k++)
                {
                    size_t slot = ((size_t)y * scn->x_row + x) * ns_ + k;
                    float v[3] = {(float)scn->ptr_r[slot], (float)scn->ptr_g[slot],
                                  (float)scn->ptr_b[slot]};
                    fwrite(v, 4, 3, fp);
                }
            }
        }
        fclose(fp);
    }

    rt_ui32 *frame = scn->get_frame();
    int fx_row = scn->get_x_row();
    if (out != nullptr)
    {
        FILE *fp = fopen(out, "wb");
        if (fp == nullptr) { perror(out); return 1; }
        for (int y = 0; y < y_res; y++)
        {
            fwrite(frame + (ptrdiff_t)y * fx_row, 4, x_res, fp);
        }
        fclose(fp);
    }
    int ns = fsaa == 2 ? 4 : fsaa == 1 ? 2 : 1;
    printf("{\"config\": %d, \"seed\": %llu, \"x_res\": %d, \"y_res\": %d, "
           "\"fsaa\": %d, \"depth\": %d, \"threads\": %d, \"simd\": %d, "
           "\"render_only\": %d, "
           "\"frames\": %d, \"median_s\": %.6f, \"total_s\": %.6f, "
           "\"msamples_per_s\": %.3f}\n",
           config, seed, x_res, y_res, fsaa, depth, g_threads, simd, render_only,
           (int)times.size(), med, sum,
           med > 0 ? (double)x_res * y_res * ns / med * 1e-6 : 0.0);
    delete scn;
    delete pfm;
    delete sc;
    return 0;
}
