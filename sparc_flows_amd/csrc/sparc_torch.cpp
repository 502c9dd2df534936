// sparc_torch.cpp -- thin PyTorch-ROCm binding of the C ABI (include/sparc_hip.h).
//
// Registers torch.classes.sparc.Solver: one sparc_handle, with float64 device
// tensors in and out.  Shapes and dtypes are checked here; all work is done
// by libsparc_hip.so on the stream given by set_stream (the Python layer
// passes torch.cuda.current_stream()).
#include <ATen/ATen.h>
#include <torch/custom_class.h>
#include <torch/library.h>

#include <string>
#include <tuple>
#include <vector>

#include "../../include/sparc_hip.h"

namespace {

void check(int rc, const sparc_handle *h, const char *what)
{
    TORCH_CHECK(rc == SPARC_OK, what, " failed (", rc, "): ", sparc_last_error(h));
}

void check_dev(const at::Tensor &t, const char *name)
{
    TORCH_CHECK(t.is_cuda(), name, " must be a GPU tensor");
    TORCH_CHECK(t.scalar_type() == at::kDouble, name, " must be float64");
    TORCH_CHECK(t.is_contiguous(), name, " must be contiguous");
}

at::Tensor cpu_f64(const at::Tensor &t, int64_t numel, const char *name)
{
    at::Tensor c = t.to(at::kCPU, at::kDouble).contiguous();
    TORCH_CHECK(c.numel() == numel, name, " has ", c.numel(), " elements, expected ", numel);
    return c;
}

struct Solver : torch::CustomClassHolder {
    sparc_handle *h = nullptr;
    int64_t nx = 0, nz = 0, nsrc = 0, adjoint = 0, device = 0;

    // model: [nz][6] (cpu or gpu); c2d, psi: [nz][nx] or empty tensors
    Solver(int64_t nx_, int64_t nz_, int64_t npmlbot, int64_t npmltop, int64_t npmlhor, int64_t nsrc_,
           int64_t adjoint_, std::vector<double> scalars, at::Tensor model, at::Tensor c2d, at::Tensor psi,
           int64_t device_)
    {
        TORCH_CHECK(scalars.size() == 8, "scalars: xlength, timestep, cadforcing, outputcad, obsheight, "
                                         "excitdep, nupeak, nuwidth");
        nx = nx_; nz = nz_; nsrc = nsrc_; adjoint = adjoint_; device = device_;
        at::Tensor m = cpu_f64(model, nz * 6, "model");
        at::Tensor c = c2d.numel() ? cpu_f64(c2d, nx * nz, "c2d") : at::Tensor();
        at::Tensor p = psi.numel() ? cpu_f64(psi, nx * nz, "psi") : at::Tensor();
        sparc_config cfg{};
        cfg.nx = (int)nx; cfg.nz = (int)nz;
        cfg.npmlbot = (int)npmlbot; cfg.npmltop = (int)npmltop; cfg.npmlhor = (int)npmlhor;
        cfg.nsrc = (int)nsrc; cfg.adjoint = (int)adjoint;
        cfg.xlength = scalars[0]; cfg.timestep = scalars[1]; cfg.cadforcing = scalars[2];
        cfg.outputcad = scalars[3]; cfg.obsheight = scalars[4]; cfg.excitdep = scalars[5];
        cfg.nupeak = scalars[6]; cfg.nuwidth = scalars[7];
        cfg.model = m.data_ptr<double>();
        cfg.c2d = c.defined() ? c.data_ptr<double>() : nullptr;
        cfg.psi = p.defined() ? p.data_ptr<double>() : nullptr;
        check(sparc_create(&cfg, &h), nullptr, "sparc_create");
    }
    ~Solver() override
    {
        if (h) sparc_destroy(h);
    }
    at::TensorOptions opts() const { return at::TensorOptions().dtype(at::kDouble).device(at::kCUDA, device); }

    void set_stream(int64_t stream) { check(sparc_set_stream(h, (void *)stream), h, "sparc_set_stream"); }

    int64_t query_int(std::string name)
    {
        int v = 0;
        check(sparc_query_int(h, name.c_str(), &v), h, "sparc_query_int");
        return v;
    }
    double query_double(std::string name)
    {
        double v = 0;
        check(sparc_query_double(h, name.c_str(), &v), h, "sparc_query_double");
        return v;
    }
    at::Tensor background(std::string name)
    {
        int n = 0;
        check(sparc_get_background(h, name.c_str(), nullptr, 0, &n), h, "sparc_get_background");
        at::Tensor out = at::empty({n}, at::TensorOptions().dtype(at::kDouble));
        check(sparc_get_background(h, name.c_str(), out.data_ptr<double>(), n, &n), h, "sparc_get_background");
        return out;
    }
    void set_model(at::Tensor c2d, at::Tensor psi)
    {
        at::Tensor c = c2d.numel() ? cpu_f64(c2d, nx * nz, "c2d") : at::Tensor();
        at::Tensor p = psi.numel() ? cpu_f64(psi, nx * nz, "psi") : at::Tensor();
        check(sparc_set_model(h, c.defined() ? c.data_ptr<double>() : nullptr,
                              p.defined() ? p.data_ptr<double>() : nullptr),
              h, "sparc_set_model");
    }
    void set_flow(at::Tensor vx, at::Tensor vz)
    {
        at::Tensor a = vx.numel() ? cpu_f64(vx, nx * nz, "vx") : at::Tensor();
        at::Tensor b = vz.numel() ? cpu_f64(vz, nx * nz, "vz") : at::Tensor();
        check(sparc_set_flow(h, a.defined() ? a.data_ptr<double>() : nullptr,
                             b.defined() ? b.data_ptr<double>() : nullptr),
              h, "sparc_set_flow");
    }
    void set_forward_sources(at::Tensor xsrc, int64_t maxtime)
    {
        at::Tensor x = cpu_f64(xsrc, nsrc, "xsrc");
        check(sparc_set_forward_sources(h, x.data_ptr<double>(), (int)maxtime), h, "sparc_set_forward_sources");
    }
    void set_forward_profile(at::Tensor vr, at::Tensor zprof, at::Tensor series, int64_t maxtime)
    {
        at::Tensor v = cpu_f64(vr, nsrc * nx, "vr");
        at::Tensor z = cpu_f64(zprof, nz, "zprof");
        TORCH_CHECK(series.numel() > 0, "series must not be empty");
        at::Tensor t = cpu_f64(series, series.numel(), "series");
        check(sparc_set_forward_profile(h, v.data_ptr<double>(), z.data_ptr<double>(), t.data_ptr<double>(),
                                        (int)t.numel(), (int)maxtime),
              h, "sparc_set_forward_profile");
    }
    void set_adjoint_sources(at::Tensor vr, int64_t maxtime)
    {
        check_dev(vr, "vr");
        TORCH_CHECK(vr.dim() == 3 && vr.size(0) == nsrc && vr.size(2) == nx, "vr must be [nsrc][nt][nx]");
        check(sparc_set_adjoint_sources(h, vr.data_ptr<double>(), (int)vr.size(1), (int)maxtime), h,
              "sparc_set_adjoint_sources");
    }
    int64_t steps() { return query_int("steps"); }

    // returns (vz [nsrc][nout][nx], snaps [nsrc][nout][nf][nz][nx] or empty)
    std::tuple<at::Tensor, at::Tensor> run(int64_t maxtime, bool snapshots)
    {
        const int64_t nout = maxtime / steps() + 1;
        at::Tensor vz = at::zeros({nsrc, nout, nx}, opts());
        at::Tensor sn = snapshots ? at::zeros({nsrc, nout, adjoint ? 4 : 6, nz, nx}, opts())
                                  : at::empty({0}, opts());
        check(sparc_run(h, (int)maxtime, vz.data_ptr<double>(), snapshots ? sn.data_ptr<double>() : nullptr),
              h, "sparc_run");
        return {vz, sn};
    }
    // steps t0..t1 into caller-allocated outputs: vz [nsrc][nslots][nx],
    // snaps [nsrc][nslots][nf][nz][nx] (either may be empty), slot j = output k0 + j
    void run_range(int64_t t0, int64_t t1, at::Tensor vz, at::Tensor sn, int64_t k0)
    {
        int64_t nslots = -1;
        if (vz.numel()) {
            check_dev(vz, "vz");
            TORCH_CHECK(vz.dim() == 3 && vz.size(0) == nsrc && vz.size(2) == nx, "vz must be [nsrc][nslots][nx]");
            nslots = vz.size(1);
        }
        if (sn.numel()) {
            check_dev(sn, "snaps");
            TORCH_CHECK(sn.dim() == 5 && sn.size(0) == nsrc && sn.size(2) == (adjoint ? 4 : 6) &&
                            sn.size(3) == nz && sn.size(4) == nx, "snaps must be [nsrc][nslots][nf][nz][nx]");
            TORCH_CHECK(nslots < 0 || nslots == sn.size(1), "vz and snaps must have the same slot count");
            nslots = sn.size(1);
        }
        check(sparc_run_range(h, (int)t0, (int)t1, vz.numel() ? vz.data_ptr<double>() : nullptr,
                              sn.numel() ? sn.data_ptr<double>() : nullptr, (int)k0, (int)std::max<int64_t>(nslots, 0)),
              h, "sparc_run_range");
    }
    int64_t checkpoint_numel()
    {
        size_t b = 0;
        check(sparc_checkpoint_size(h, &b), h, "sparc_checkpoint_size");
        return (int64_t)(b / sizeof(double));
    }
    void save_checkpoint(at::Tensor buf)
    {
        check_dev(buf, "checkpoint");
        TORCH_CHECK(buf.numel() >= checkpoint_numel(), "checkpoint buffer too small");
        check(sparc_save_checkpoint(h, buf.data_ptr<double>()), h, "sparc_save_checkpoint");
    }
    void load_checkpoint(at::Tensor buf)
    {
        check_dev(buf, "checkpoint");
        TORCH_CHECK(buf.numel() >= checkpoint_numel(), "checkpoint buffer too small");
        check(sparc_load_checkpoint(h, buf.data_ptr<double>()), h, "sparc_load_checkpoint");
    }
    void kernels_begin() { check(sparc_kernels_begin(h), h, "sparc_kernels_begin"); }
    void kernels_add(at::Tensor fwd, at::Tensor adj)
    {
        check_dev(fwd, "fwd");
        check_dev(adj, "adj");
        TORCH_CHECK(fwd.dim() == 5 && fwd.size(0) == nsrc && fwd.size(2) == 6 && fwd.size(3) == nz &&
                        fwd.size(4) == nx, "fwd must be [nsrc][n][6][nz][nx]");
        TORCH_CHECK(adj.dim() == 5 && adj.size(0) == nsrc && adj.size(1) == fwd.size(1) && adj.size(2) == 4,
                    "adj must be [nsrc][n][4][nz][nx]");
        check(sparc_kernels_add(h, (int)fwd.size(1), fwd.data_ptr<double>(), adj.data_ptr<double>()), h,
              "sparc_kernels_add");
    }
    std::vector<at::Tensor> kernels_end()
    {
        std::vector<at::Tensor> o;
        for (int i = 0; i < 5; i++) o.push_back(at::empty({nsrc, nz, nx}, opts()));
        check(sparc_kernels_end(h, o[0].data_ptr<double>(), o[1].data_ptr<double>(), o[2].data_ptr<double>(),
                                o[3].data_ptr<double>(), o[4].data_ptr<double>()),
              h, "sparc_kernels_end");
        return o;
    }
    void reset() { check(sparc_reset(h), h, "sparc_reset"); }
    void step(int64_t t) { check(sparc_step(h, (int)t), h, "sparc_step"); }
    at::Tensor get_state()
    {
        at::Tensor a = at::empty({4, nsrc, nz, nx}, opts());
        check(sparc_get_state(h, a.data_ptr<double>()), h, "sparc_get_state");
        return a;
    }
    void set_state(at::Tensor a)
    {
        check_dev(a, "state");
        TORCH_CHECK(a.numel() == 4 * nsrc * nz * nx, "state must be [4][nsrc][nz][nx]");
        check(sparc_set_state(h, a.data_ptr<double>()), h, "sparc_set_state");
    }
    std::tuple<at::Tensor, at::Tensor> get_pml()
    {
        const int64_t nzpml = query_int("nzpml");
        const int64_t nph = query_int("nph");
        at::Tensor pv = at::empty({2, nsrc, nzpml, nx}, opts());
        at::Tensor px = at::empty({2, nsrc, nz, nph}, opts());
        check(sparc_get_pml(h, pv.data_ptr<double>(), px.data_ptr<double>()), h, "sparc_get_pml");
        return {pv, px};
    }
    at::Tensor ddx(at::Tensor in)
    {
        check_dev(in, "in");
        TORCH_CHECK(in.numel() % nx == 0 && in.size(-1) == nx, "last dimension must be nx");
        at::Tensor out = at::empty_like(in);
        check(sparc_ddx(h, in.data_ptr<double>(), out.data_ptr<double>(), (int)(in.numel() / nx)), h, "sparc_ddx");
        return out;
    }
    at::Tensor ddz_impl(at::Tensor in, int64_t ibc, at::Tensor preset, bool kern)
    {
        check_dev(in, "in");
        TORCH_CHECK(in.dim() >= 2 && in.size(-1) == nx && in.size(-2) == nz, "trailing dims must be [nz][nx]");
        at::Tensor out = preset.numel() ? preset.to(in.device()).clone().contiguous() : at::zeros_like(in);
        TORCH_CHECK(out.sizes() == in.sizes(), "preset must match the input shape");
        const int nb = (int)(in.numel() / (nx * nz));
        const int rc = kern ? sparc_ddz_kern(h, in.data_ptr<double>(), out.data_ptr<double>(), nb, (int)ibc)
                            : sparc_ddz(h, in.data_ptr<double>(), out.data_ptr<double>(), nb, (int)ibc);
        check(rc, h, "sparc_ddz");
        return out;
    }
    at::Tensor ddz(at::Tensor in, int64_t ibc, at::Tensor preset) { return ddz_impl(in, ibc, preset, false); }
    at::Tensor ddz_kern(at::Tensor in, int64_t ibc) { return ddz_impl(in, ibc, at::Tensor(), true); }

    // per-step sum of v_z^2 into norm2 [nsrc][n] from time t0 on (empty: off)
    void set_norm_output(at::Tensor norm2, int64_t t0)
    {
        if (!norm2.numel()) {
            check(sparc_set_norm_output(h, nullptr, 0, 0), h, "sparc_set_norm_output");
            return;
        }
        check_dev(norm2, "norm2");
        TORCH_CHECK(norm2.dim() == 2 && norm2.size(0) == nsrc, "norm2 must be [nsrc][n]");
        check(sparc_set_norm_output(h, norm2.data_ptr<double>(), (int)t0, (int)norm2.size(1)), h,
              "sparc_set_norm_output");
    }

    // exact discrete adjoint: gvz [nsrc][nt][nx] -> gpsi, gc, gcoef ([3][nsrc][nz][nx])
    std::vector<at::Tensor> discrete_gradient(int64_t maxtime, at::Tensor gvz, int64_t interval, bool want_psi)
    {
        check_dev(gvz, "gvz");
        TORCH_CHECK(gvz.dim() == 3 && gvz.size(0) == nsrc && gvz.size(2) == nx, "gvz must be [nsrc][nt][nx]");
        at::Tensor gpsi = at::empty({nsrc, nz, nx}, opts()), gc = at::empty({nsrc, nz, nx}, opts());
        at::Tensor gcoef = at::empty({3, nsrc, nz, nx}, opts());
        check(sparc_discrete_gradient(h, (int)maxtime, gvz.data_ptr<double>(), (int)gvz.size(1), (int)interval,
                                      want_psi ? gpsi.data_ptr<double>() : nullptr, gc.data_ptr<double>(),
                                      gcoef.data_ptr<double>()),
              h, "sparc_discrete_gradient");
        return {gpsi, gc, gcoef};
    }
    // transpose of ddz (ibc 1..3); returns out and, for ibc 3, the adjoint of the kept last row
    std::vector<at::Tensor> ddz_T(at::Tensor in, int64_t ibc)
    {
        check_dev(in, "in");
        TORCH_CHECK(in.numel() % (nx * nz) == 0, "in must hold whole [nz][nx] fields");
        const int64_t nb = in.numel() / (nx * nz);
        at::Tensor out = at::empty_like(in), pb = at::zeros({nb, nx}, opts());
        check(sparc_ddz_T(h, in.data_ptr<double>(), out.data_ptr<double>(), (int)nb, (int)ibc, pb.data_ptr<double>()),
              h, "sparc_ddz_T");
        return {out, pb};
    }

    std::vector<double> profile_kernels(int64_t reps)
    {
        std::vector<double> ms(5, 0.0);
        check(sparc_profile_kernels(h, (int)reps, ms.data()), h, "sparc_profile_kernels");
        return ms;
    }

    // returns kc, kpsi, kvx, kvz, hess, each [nsrc][nz][nx]
    std::vector<at::Tensor> kernels(at::Tensor fwd, at::Tensor adj)
    {
        check_dev(fwd, "fwd");
        check_dev(adj, "adj");
        TORCH_CHECK(fwd.dim() == 5 && fwd.size(0) == nsrc && fwd.size(2) == 6 && fwd.size(3) == nz &&
                        fwd.size(4) == nx, "fwd must be [nsrc][nt][6][nz][nx]");
        TORCH_CHECK(adj.dim() == 5 && adj.size(0) == nsrc && adj.size(1) == fwd.size(1) && adj.size(2) == 4,
                    "adj must be [nsrc][nt][4][nz][nx]");
        std::vector<at::Tensor> o;
        for (int i = 0; i < 5; i++) o.push_back(at::empty({nsrc, nz, nx}, opts()));
        check(sparc_kernels(h, (int)fwd.size(1), fwd.data_ptr<double>(), adj.data_ptr<double>(),
                            o[0].data_ptr<double>(), o[1].data_ptr<double>(), o[2].data_ptr<double>(),
                            o[3].data_ptr<double>(), o[4].data_ptr<double>()),
              h, "sparc_kernels");
        return o;
    }
};

}  // namespace

TORCH_LIBRARY(sparc, m)
{
    m.class_<Solver>("Solver")
        .def(torch::init<int64_t, int64_t, int64_t, int64_t, int64_t, int64_t, int64_t, std::vector<double>,
                         at::Tensor, at::Tensor, at::Tensor, int64_t>())
        .def("set_stream", &Solver::set_stream)
        .def("query_int", &Solver::query_int)
        .def("query_double", &Solver::query_double)
        .def("background", &Solver::background)
        .def("set_model", &Solver::set_model)
        .def("set_flow", &Solver::set_flow)
        .def("run_range", &Solver::run_range)
        .def("checkpoint_numel", &Solver::checkpoint_numel)
        .def("save_checkpoint", &Solver::save_checkpoint)
        .def("load_checkpoint", &Solver::load_checkpoint)
        .def("kernels_begin", &Solver::kernels_begin)
        .def("kernels_add", &Solver::kernels_add)
        .def("kernels_end", &Solver::kernels_end)
        .def("set_forward_sources", &Solver::set_forward_sources)
        .def("set_forward_profile", &Solver::set_forward_profile)
        .def("set_adjoint_sources", &Solver::set_adjoint_sources)
        .def("run", &Solver::run)
        .def("reset", &Solver::reset)
        .def("step", &Solver::step)
        .def("get_state", &Solver::get_state)
        .def("set_state", &Solver::set_state)
        .def("get_pml", &Solver::get_pml)
        .def("ddx", &Solver::ddx)
        .def("ddz", &Solver::ddz)
        .def("ddz_kern", &Solver::ddz_kern)
        .def("kernels", &Solver::kernels)
        .def("profile_kernels", &Solver::profile_kernels)
        .def("discrete_gradient", &Solver::discrete_gradient)
        .def("set_norm_output", &Solver::set_norm_output)
        .def("ddz_T", &Solver::ddz_T);
}
