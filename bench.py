"""Benchmark: airfoil finite-volume time steps on one MI355X.

Workload: the 1200 x 600 synthetic quad mesh (720,000 cells, 1,438,200
interior edges, 3,600 boundary edges, numbering permuted).  A step is one
airfoil iteration: save_soln, then 2 x (adt_calc, res_calc, bres_calc,
update with the rms Global).  ``value`` is algorithmic (unique) bytes per
step x steps / wall time of the timed region, inputs resident in HBM.

``roofline`` is for res_calc, the dominant loop: its unique bytes per call
divided by its average duration, timed with HIP events on the stream the
loop is launched on.  ``cache_cold`` times res_calc and adt_calc each
after a 512 MiB write and read, so that their inputs (about 110 and 52 MB, which
would fit the 256 MB Infinity Cache) come from HBM.  ``cpu_baseline`` times the reference's OpenMP backend
restated (block colouring, the host's OpenMP threads), with the sequential
oracle (one core) beside it, on a bounded number of steps of the same
mesh.

Prints one JSON line.  ``--gpus N`` with N > 1 exits: this build measures
one GPU.
"""

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# bytes written between the launches of the cache-cold loop timings
FLUSH_BYTES = 512 << 20
# GPU clock cycles of torch.cuda._sleep before a singly timed launch (about
# 110 us on MI355X: longer than the host takes to enqueue one par_loop)
LAUNCH_DELAY_CYCLES = 1000


def parse():
    p = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=1000)
    p.add_argument('--warmup', type=int, default=20)
    p.add_argument('--nx', type=int, default=1200)
    p.add_argument('--vector-maps', type=int, default=0,
                   help='drive the step in the vector-map form (1) instead of the scalar one')
    p.add_argument('--ny', type=int, default=600)
    p.add_argument('--soa', type=int, default=0,
                   help='store q, q_old and res as structure-of-arrays Dats (1)')
    p.add_argument('--graph', type=int, default=0,
                   help='replay each step as a hipGraph (default: par_loops called eagerly)')
    p.add_argument('--cold', type=int, default=1,
                   help='time res_calc and adt_calc cache-cold too (0: skip)')
    p.add_argument('--cpu-seconds', type=float, default=15.0,
                   help='approximate CPU time budget of the oracle baseline (0: skip)')
    p.add_argument('--partition-size', type=int, default=0)
    p.add_argument('--block-size', type=int, default=0)
    p.add_argument('--autotune', type=int, default=0,
                   help='tune launch parameters per loop and mesh-size bucket (1: on)')
    p.add_argument('--owner-stage', type=int, default=-1,
                   help='owner blocks stage owned rows of gathered Dats in LDS (-1: default)')
    p.add_argument('--owner-block-size', type=int, default=0)
    p.add_argument('--owner-fill', type=float, default=0,
                   help='elements per thread owner blocks are sized for (0: the default)')
    p.add_argument('--owner-prefetch', type=int, default=-1,
                   help='owner write-back loads issued before the elements (-1: default)')
    p.add_argument('--owner-waves', type=int, default=-1,
                   help='waves per SIMD owner loops are compiled for (-1: default, 0: none)')
    p.add_argument('--stream-unroll', type=int, default=0,
                   help='elements per thread of direct loops (0: the default)')
    p.add_argument('--stream-blocks', type=int, default=0,
                   help='most workgroups of a direct loop (0: the default)')
    p.add_argument('--layout-order', default='', help="'bisect' or 'bfs' (default: the backend's)")
    p.add_argument('--stage-reads', type=int, default=-1,
                   help='stage read-only multi-component direct Dats through LDS (-1: default)')
    p.add_argument('--stage-writes', type=int, default=-1,
                   help='stage written multi-component direct Dats through LDS (-1: default)')
    p.add_argument('--owner-targets', type=int, default=0,
                   help='targets owned per block of owner-computes loops (0: sized automatically)')
    p.add_argument('--assembly', type=int, default=128,
                   help='edge count n of the n^3 Kuhn-tet assembly measurement (0: skip)')
    p.add_argument('--batch', type=int, default=256,
                   help='meshes in the batched assembly+CG measurement (0: skip)')
    p.add_argument('--batch-emin', type=float, default=2e3)
    p.add_argument('--batch-emax', type=float, default=2e6)
    p.add_argument('--batch-rounds', type=int, default=1,
                   help='timed rounds of each batching path (after one untimed round)')
    p.add_argument('--serve', type=int, default=64,
                   help='requests per arrival rate in the serving measurement (0: skip)')
    p.add_argument('--serve-pool', type=int, default=32,
                   help='2-D and 3-D meshes each in the serving pool')
    p.add_argument('--serve-tri', type=float, nargs=2, default=(8e3, 5e5),
                   help='element range of the pool\'s triangle meshes')
    p.add_argument('--serve-tet', type=float, nargs=2, default=(5e4, 2e6),
                   help='element range of the pool\'s tetrahedral meshes')
    p.add_argument('--solvers', type=float, default=2e6,
                   help='elements of the 2-D system the CG, GMRES(30) and BiCGStab '
                        'iterations are timed on (0: skip)')
    p.add_argument('--opt', action='append', default=[], metavar='KEY=VALUE',
                   help='set a backend configuration key (repeatable), e.g. '
                        'hip_owner_mat_block_size=512')
    p.add_argument('--copy-calibration', action='store_true',
                   help='only run a 256 MiB device copy three times (the reference reading '
                        'for rocprofv3 FETCH_SIZE/WRITE_SIZE) and exit')
    return p.parse_args()


def main():
    args = parse()
    if args.copy_calibration:
        import torch
        n = 256 * 1024 * 1024 // 8
        a = torch.ones(n, dtype=torch.float64, device='cuda')
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({'copy_bytes': n * 8, 'copies': 3}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU; --gpus %d is not supported (every run uses one "
              "MI355X)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from pyop2_amd import op2, hip
    from pyop2_amd import _native as nat
    from pyop2_amd.workloads import airfoil, meshes

    opts = {'backend': 'hip'}
    if args.partition_size:
        opts['hip_partition_size'] = args.partition_size
    if args.block_size:
        opts['hip_block_size'] = args.block_size
    if args.stage_reads >= 0:
        opts['hip_stage_reads'] = bool(args.stage_reads)
    if args.stage_writes >= 0:
        opts['hip_stage_writes'] = bool(args.stage_writes)
    if args.layout_order:
        opts['hip_layout_order'] = args.layout_order
    if args.owner_prefetch >= 0:
        opts['hip_owner_prefetch'] = args.owner_prefetch
    if args.owner_waves >= 0:
        opts['hip_owner_waves'] = args.owner_waves
    if args.owner_stage >= 0:
        opts['hip_owner_stage'] = bool(args.owner_stage)
    if args.autotune:
        opts['hip_autotune'] = True
    if args.stream_unroll:
        opts['hip_stream_unroll'] = args.stream_unroll
    if args.stream_blocks:
        opts['hip_stream_blocks'] = args.stream_blocks
    if args.owner_block_size:
        opts['hip_owner_block_size'] = args.owner_block_size
    if args.owner_fill:
        opts['hip_owner_fill'] = args.owner_fill
    if args.owner_targets:
        opts['hip_owner_targets'] = args.owner_targets
    for kv in args.opt:
        k, v = kv.split('=', 1)
        try:
            v = json.loads(v)
        except ValueError:
            pass
        opts[k] = v
    op2.init(**opts)
    mesh = meshes.QuadMesh(args.nx, args.ny, seed=0)
    A = (airfoil.AirfoilVector if args.vector_maps else airfoil.Airfoil)(op2, mesh,
                                                                         soa=bool(args.soa))
    ub = airfoil.unique_bytes(mesh)
    step_bytes = ub['save_soln'] + 2 * (ub['adt_calc'] + ub['res_calc'] + ub['bres_calc'] +
                                        ub['update'])
    s = hip.stream()
    # a captured step needs one eager step before it (compilation, plans,
    # uploads)
    for _ in range(max(args.warmup, 1 if args.graph else 0)):
        A.iteration()
    torch.cuda.synchronize()

    step = A.iteration
    graph = None
    if args.graph:
        graph = hip.capture(A.iteration)
        graph.replay()
        torch.cuda.synchronize()
        step = graph.replay

    def mkev():
        e = ctypes.c_void_p()
        nat.check(nat.lib.op2hip_event_create(ctypes.byref(e)))
        return e

    t_start, t_stop = mkev(), mkev()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nat.lib.op2hip_event_record(t_start, ctypes.c_void_p(s))
    for _ in range(args.steps):
        step()
    # the last step's deferred Global reduction runs inside the timed region
    hip.flush()
    nat.lib.op2hip_event_record(t_stop, ctypes.c_void_p(s))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ctypes.c_float()
    nat.lib.op2hip_event_elapsed_ms(t_start, t_stop, ctypes.byref(ms))
    gpu_s = ms.value / 1e3
    elapsed = max(wall, gpu_s)

    def hot_avg(loop, nrep=10):
        # one loop alone: a graph of `nrep` back-to-back calls, HIP events
        # on the stream around its replays (no host time in between)
        g = hip.capture(lambda: [loop() for _ in range(nrep)])
        g.replay()
        torch.cuda.synchronize()
        a, b = mkev(), mkev()
        reps = max(3, args.steps // 5)
        nat.lib.op2hip_event_record(a, ctypes.c_void_p(s))
        for _ in range(reps):
            g.replay()
        nat.lib.op2hip_event_record(b, ctypes.c_void_p(s))
        torch.cuda.synchronize()
        ms = ctypes.c_float()
        nat.lib.op2hip_event_elapsed_ms(a, b, ctypes.byref(ms))
        return ms.value / 1e3 / (reps * nrep)

    def single_avg(loop, flush=None, nrep=40):
        # one loop launched on its own, eagerly, behind a GPU-side delay
        # long enough for the host to have enqueued the launch before it
        # ends, so the events time the kernel and its dispatch (not the
        # host's launch, nor a graph launch's latency); with `flush`, after
        # a write and then a read of `flush` (larger than the L2s and the
        # 256 MB Infinity Cache together), so its inputs come from HBM (the
        # read leaves the caches holding clean lines, so the loop does not
        # pay for writing back the flush's own data)
        loop()
        torch.cuda.synchronize()
        evs = [(mkev(), mkev()) for _ in range(nrep)]
        for i, (a, b) in enumerate(evs):
            if flush is not None:
                flush.fill_(float(i))
                flush.sum()
            torch.cuda._sleep(LAUNCH_DELAY_CYCLES)
            nat.lib.op2hip_event_record(a, ctypes.c_void_p(s))
            loop()
            nat.lib.op2hip_event_record(b, ctypes.c_void_p(s))
        torch.cuda.synchronize()
        tot, ms = 0.0, ctypes.c_float()
        for a, b in evs:
            nat.lib.op2hip_event_elapsed_ms(a, b, ctypes.byref(ms))
            tot += ms.value
        return tot / 1e3 / nrep

    # the step as a program written for the reference drives it: each
    # par_loop called from Python (no graph)
    eager_steps = max(20, args.steps // 5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(eager_steps):
        A.iteration()
    hip.flush()
    torch.cuda.synchronize()
    eager_s = (time.perf_counter() - t0) / eager_steps
    # as the reference's demo writes it: a new rms Global per update, read
    # back on the host after every step
    A.demo_iteration()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(eager_steps):
        A.demo_iteration()
    demo_s = (time.perf_counter() - t0) / eager_steps

    res_avg_s = hot_avg(A.res_calc)
    res_gbs = ub['res_calc'] / res_avg_s / 1e9
    cold = None
    if args.cold:
        flush = torch.empty(FLUSH_BYTES // 8, dtype=torch.float64, device='cuda')
        cold = {'flush_bytes': FLUSH_BYTES}
        for name, loop in (('res_calc', A.res_calc), ('adt_calc', A.adt_calc)):
            hot_s = res_avg_s if name == 'res_calc' else hot_avg(loop)
            single_s = single_avg(loop)
            cold_s = single_avg(loop, flush)
            cold[name] = {'unique_bytes': int(ub[name]), 'hot_us': round(hot_s * 1e6, 2),
                          'hot_frac': round(ub[name] / hot_s / 1e9 / HBM_PEAK_GBS, 4),
                          'single_us': round(single_s * 1e6, 2),
                          'cold_us': round(cold_s * 1e6, 2),
                          'cold_frac': round(ub[name] / cold_s / 1e9 / HBM_PEAK_GBS, 4)}
        del flush

    value = step_bytes * args.steps / elapsed / 1e9
    out = {
        # the headline metric: value is its first part (par_loop GB/s over
        # the airfoil step); the assembly and batching/serving parts are in
        # extra (nnz/s; solves/s with p50/p99)
        'metric': 'par_loop HBM GB/s (% of peak); CSR nnz/s assembled; batched solves/s + p99',
        'value': round(value, 2),
        'unit': 'GB/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(elapsed / args.steps * 1e3, 4),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic',
        'config': {'workload': 'airfoil %dx%d quad mesh, permuted numbering, '
                               'save_soln + 2x(adt_calc, res_calc, bres_calc, update) per step'
                               % (args.nx, args.ny),
                   'cells': mesh.ncell, 'edges': mesh.nedge, 'bedges': mesh.nbedge,
                   'unique_bytes_per_step': int(step_bytes),
                   'pct_of_hbm_peak': round(100.0 * value / HBM_PEAK_GBS, 2),
                   'graph': bool(args.graph),
                   'eager_ms_per_step': round(eager_s * 1e3, 4),
                   'demo_ms_per_step': round(demo_s * 1e3, 4),
                   'vector_maps': bool(args.vector_maps),
                   'soa': bool(args.soa),
                   'autotune': _tuned() if args.autotune else False,
                   'partition_size': hip.cfg.get('hip_partition_size') or 'auto'},
        'roofline': {'bound': 'hbm', 'kernel': 'op2_res_calc_owner (one res_calc par_loop)',
                     'achieved': round(res_gbs, 2), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': round(res_gbs / HBM_PEAK_GBS, 4),
                     'avg_us': round(res_avg_s * 1e6, 2),
                     'bytes_per_launch': int(ub['res_calc']), 'traffic': None},
        # res_calc and adt_calc alone, back to back (hot) and each after a
        # 512 MiB write and read that evict L2 and the Infinity Cache (cold)
        'cache_cold': cold,
    }
    if args.cpu_seconds > 0:
        out['cpu_baseline'] = cpu_baseline(mesh, args.cpu_seconds, step_bytes)
    A.remove_consts()
    traffic = _profiled_traffic()
    if traffic is not None:
        out['roofline']['traffic'] = traffic
    extra = {}
    if args.assembly:
        extra['assembly'] = assembly(op2, args.assembly)
    if args.batch:
        from pyop2_amd.workloads import suites
        extra['batched'] = suites.batch_suite(op2, nmesh=args.batch, emin=args.batch_emin,
                                              emax=args.batch_emax, rounds=args.batch_rounds,
                                              progress=_progress('batch'))
    if args.serve:
        from pyop2_amd.workloads import suites
        pool = suites.make_pool(op2, n2d=args.serve_pool, n3d=args.serve_pool,
                                tri_range=tuple(args.serve_tri),
                                tet_range=tuple(args.serve_tet))
        _progress('serve')('pool of %d meshes resident' % len(pool))
        extra['serving'] = suites.serve_suite(op2, pool, nreq=args.serve,
                                              progress=_progress('serve'))
    if args.solvers:
        from pyop2_amd.workloads import cgbench
        extra['solvers'] = cgbench.run(op2, int(args.solvers), 200)
    if extra:
        out['extra'] = extra
    if args.autotune:
        # every leg's loops, tuned as they first ran
        out['config']['autotune'] = _tuned()
    print(json.dumps(out))


def _progress(leg):
    """A progress callback writing to stderr (stdout carries only the JSON
    line)."""
    t0 = time.perf_counter()

    def say(msg):
        sys.stderr.write('[bench %s %.1fs] %s\n' % (leg, time.perf_counter() - t0, msg))
        sys.stderr.flush()
    return say


def _tuned():
    """Parameters the autotuner chose, per loop signature."""
    from pyop2_amd import autotune
    return {k.split('/')[0] + ' ' + k.split('/')[-1]: v['params']
            for k, v in autotune.table().items()}


def _profiled_traffic():
    """HBM bytes per res_calc loop measured by rocprofv3 FETCH_SIZE and
    WRITE_SIZE (calibrated against a copy kernel), when recorded."""
    path = os.path.join(ROOT, 'profiles', 'res_calc_airfoil_1200x600_traffic.json')
    try:
        with open(path) as f:
            return json.load(f).get('traffic_bytes_per_loop')
    except (IOError, ValueError):
        return None


def assembly(op2, n):
    """CSR nonzeros assembled per second on the n^3 Kuhn-tet mesh: P1
    Laplace with a lognormal coefficient and vector-P1 mass."""
    import numpy as np
    import torch
    from pyop2_amd.workloads import fem, meshes
    t = time.perf_counter()
    mesh = meshes.TetMesh(n, seed=0)
    kappa = np.random.default_rng(1).lognormal(0.0, 1.0, mesh.nvert)
    t_mesh = time.perf_counter() - t
    t = time.perf_counter()
    a = fem.Assembly3D(op2, mesh, kappa, name='bench3d')
    torch.cuda.synchronize()
    setup = time.perf_counter() - t
    res = {'mesh': '%d^3 Kuhn tets' % n, 'tets': mesh.nelem, 'vertices': mesh.nvert,
           'mesh_generation_s': round(t_mesh, 2),
           'layout_and_sparsity_s': round(setup, 2),
           'laplace_nnz': a.sp_scalar.nz, 'laplace_nnz_closed_form': fem.kuhn_nnz(n),
           'vector_mass_nnz': a.sp_vector.nz}
    # the device CSR build alone (maps already on the device), timed on a
    # rebuild into buffers of its own
    for name, sp in (('laplace', a.sp_scalar), ('vector_mass', a.sp_vector)):
        sp._build_device()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(3):
            out = sp._build_device()
            assert out[3] == sp._d_nz
        torch.cuda.synchronize()
        res[name + '_sparsity_build_ms'] = round((time.perf_counter() - t) / 3 * 1e3, 2)
        del out
    for name, fn, nnz in (('laplace', a.laplace, a.sp_scalar.nz),
                          ('vector_mass', a.vector_mass, a.sp_vector.nz)):
        fn()
        torch.cuda.synchronize()
        reps = 5
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / reps
        res[name + '_ms'] = round(dt * 1e3, 3)
        res[name + '_nnz_per_s'] = float('%.4g' % (nnz / dt))
    return res


def cpu_baseline(mesh, budget_s, step_bytes):
    """The CPU baselines on whole steps of the same mesh: the OpenMP
    restatement of the reference's OpenMP backend on the host's cores (the
    reported baseline) and the sequential oracle on one core, each for about
    half of ``budget_s``."""
    from oracle import openmp as omp
    from oracle import sequential as seq
    from pyop2_amd.workloads import airfoil

    def run(par_loop, budget):
        B = airfoil.Airfoil(seq, mesh, consts=False)
        if par_loop is not None:
            B.W = _World(par_loop)
        B.iteration()  # compile the wrappers, build the plans
        t0 = time.perf_counter()
        n = 0
        while True:
            B.iteration()
            n += 1
            if time.perf_counter() - t0 >= budget:
                break
        return n, time.perf_counter() - t0

    n_s, dt_s = run(None, budget_s / 2)
    n_o, dt_o = run(omp.par_loop, budget_s / 2)
    cores = omp.threads()
    return {'value': round(step_bytes * n_o / dt_o / 1e9, 3), 'unit': 'GB/s', 'cores': cores,
            'kind': 'port',
            'sample': '%d airfoil steps on the same mesh, OpenMP restatement of the reference\'s '
                      'openmp backend (block colouring, part size %d, gcc -O2 -fopenmp, %d '
                      'threads), %.1f ms/step' % (n_o, omp.PART_SIZE, cores, dt_o / n_o * 1e3),
            'sequential': {'value': round(step_bytes * n_s / dt_s / 1e9, 3), 'unit': 'GB/s',
                           'cores': 1, 'kind': 'port',
                           'sample': '%d steps, sequential oracle (gcc -O2), %.1f ms/step'
                                     % (n_s, dt_s / n_s * 1e3)}}


class _World(object):
    """The sequential oracle's namespace with another ``par_loop``."""

    def __init__(self, par_loop):
        from oracle import sequential as seq
        self.__dict__.update({k: getattr(seq, k) for k in dir(seq) if not k.startswith('__')})
        self.par_loop = par_loop


if __name__ == '__main__':
    main()
