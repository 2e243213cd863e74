// Minimal JSON reader for the solver's parameter object (no external JSON library is available
// to this build).  Supports the full JSON grammar; only objects, numbers, booleans and strings
// are interpreted.
#include "params.hpp"
#include "common.hpp"
#include "kernels.hpp"

static_assert(psmas::RESTRICT_SMALL_MAX == 8 && psmas::RESTRICT_PIECE == 512, "restrict_piece range message");

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <map>
#include <memory>
#include <sstream>
#include <vector>

namespace psmas
{
    namespace
    {
        struct JValue
        {
            enum Kind
            {
                Null,
                Bool,
                Number,
                String,
                Array,
                Object
            } kind = Null;
            bool b = false;
            double num = 0.0;
            bool is_integer = false;
            std::string str;
            std::vector<JValue> arr;
            std::vector<std::pair<std::string, JValue>> obj;

            const JValue *get(const std::string &key) const
            {
                for (const auto &kv : obj)
                    if (kv.first == key)
                        return &kv.second;
                return nullptr;
            }
        };

        class Parser
        {
        public:
            explicit Parser(const std::string &s) : s_(s) {}

            JValue parse()
            {
                JValue v = value();
                ws();
                if (i_ != s_.size())
                    fail("trailing characters");
                return v;
            }

        private:
            const std::string &s_;
            size_t i_ = 0;

            [[noreturn]] void fail(const char *what)
            {
                throw InvalidArgumentError(std::string("[MAS] invalid parameter JSON: ") + what +
                                           " at offset " + std::to_string(i_));
            }
            void ws()
            {
                while (i_ < s_.size() && std::isspace(static_cast<unsigned char>(s_[i_])))
                    ++i_;
            }
            bool eat(char c)
            {
                ws();
                if (i_ < s_.size() && s_[i_] == c)
                {
                    ++i_;
                    return true;
                }
                return false;
            }
            void expect(char c)
            {
                if (!eat(c))
                    fail("unexpected character");
            }
            JValue value()
            {
                ws();
                if (i_ >= s_.size())
                    fail("unexpected end");
                const char c = s_[i_];
                if (c == '{')
                    return object();
                if (c == '[')
                    return array();
                if (c == '"')
                {
                    JValue v;
                    v.kind = JValue::String;
                    v.str = string();
                    return v;
                }
                if (s_.compare(i_, 4, "true") == 0)
                {
                    i_ += 4;
                    JValue v;
                    v.kind = JValue::Bool;
                    v.b = true;
                    return v;
                }
                if (s_.compare(i_, 5, "false") == 0)
                {
                    i_ += 5;
                    JValue v;
                    v.kind = JValue::Bool;
                    return v;
                }
                if (s_.compare(i_, 4, "null") == 0)
                {
                    i_ += 4;
                    return JValue();
                }
                return number();
            }
            JValue object()
            {
                expect('{');
                JValue v;
                v.kind = JValue::Object;
                if (eat('}'))
                    return v;
                do
                {
                    ws();
                    std::string k = string();
                    expect(':');
                    JValue item = value();
                    v.obj.emplace_back(std::move(k), std::move(item));
                } while (eat(','));
                expect('}');
                return v;
            }
            JValue array()
            {
                expect('[');
                JValue v;
                v.kind = JValue::Array;
                if (eat(']'))
                    return v;
                do
                {
                    v.arr.push_back(value());
                } while (eat(','));
                expect(']');
                return v;
            }
            std::string string()
            {
                if (i_ >= s_.size() || s_[i_] != '"')
                    fail("expected string");
                ++i_;
                std::string out;
                while (i_ < s_.size() && s_[i_] != '"')
                {
                    char c = s_[i_++];
                    if (c == '\\')
                    {
                        if (i_ >= s_.size())
                            fail("bad escape");
                        char e = s_[i_++];
                        switch (e)
                        {
                        case 'n':
                            out.push_back('\n');
                            break;
                        case 't':
                            out.push_back('\t');
                            break;
                        case 'r':
                            out.push_back('\r');
                            break;
                        case 'b':
                            out.push_back('\b');
                            break;
                        case 'f':
                            out.push_back('\f');
                            break;
                        case 'u':
                            if (i_ + 4 > s_.size())
                                fail("bad unicode escape");
                            i_ += 4;
                            out.push_back('?');
                            break;
                        default:
                            out.push_back(e);
                        }
                    }
                    else
                    {
                        out.push_back(c);
                    }
                }
                if (i_ >= s_.size())
                    fail("unterminated string");
                ++i_;
                return out;
            }
            JValue number()
            {
                const size_t start = i_;
                if (i_ < s_.size() && (s_[i_] == '-' || s_[i_] == '+'))
                    ++i_;
                bool integer = true;
                while (i_ < s_.size() && (std::isdigit(static_cast<unsigned char>(s_[i_])) || s_[i_] == '.' ||
                                          s_[i_] == 'e' || s_[i_] == 'E' || s_[i_] == '-' || s_[i_] == '+'))
                {
                    if (s_[i_] == '.' || s_[i_] == 'e' || s_[i_] == 'E')
                        integer = false;
                    ++i_;
                }
                if (start == i_)
                    fail("unexpected token");
                const std::string tok = s_.substr(start, i_ - start);
                char *end = nullptr;
                const double d = std::strtod(tok.c_str(), &end);
                if (end == tok.c_str() || *end != '\0')
                    fail("bad number");
                JValue v;
                v.kind = JValue::Number;
                v.num = d;
                v.is_integer = integer;
                return v;
            }
        };

        [[noreturn]] void bad(const std::string &key, const char *why)
        {
            throw InvalidArgumentError("[MAS] parameter '" + key + "': " + why);
        }

        int as_int(const JValue &v, const std::string &key)
        {
            if (v.kind != JValue::Number || std::floor(v.num) != v.num)
                bad(key, "expected an integer");
            return static_cast<int>(v.num);
        }
        double as_double(const JValue &v, const std::string &key)
        {
            if (v.kind != JValue::Number)
                bad(key, "expected a number");
            if (!std::isfinite(v.num)) // e.g. 1e999: out of double range
                bad(key, "expected a finite number");
            return v.num;
        }
        bool as_bool(const JValue &v, const std::string &key)
        {
            if (v.kind != JValue::Bool)
                bad(key, "expected a boolean");
            return v.b;
        }
        std::string as_string(const JValue &v, const std::string &key)
        {
            if (v.kind != JValue::String)
                bad(key, "expected a string");
            return v.str;
        }

        const char *known_keys[] = {
            "block_dim", "max_iter", "relative_tolerance", "absolute_tolerance", "lazy_partitioning",
            "use_preconditioned_residual_norm", "precision", "aggregation", "coarse_solver",
            "reliable_update_delta", "check_interval", "true_residual_period", "use_graphs",
            "learned_correction", "division_guard", "learned_collect", "lattice_points", "learned_hidden",
            "learned_random_pairs", "learned_snapshots", "learned_stall_segments", "learned_output_scale", "spmv_unroll", "spmv_halo", "device_setup", "spmv_copies", "spmv_fuse_direction",
            "restrict_piece", "apply_workgroups", "learned_network", "apply_wave", "single_reduction", "spmv_symmetric"};
    } // namespace

    void apply_json_parameters(Params &p, const std::string &json)
    {
        const JValue root = Parser(json).parse();
        if (root.kind != JValue::Object)
            throw InvalidArgumentError("[MAS] parameters must be a JSON object");
        const JValue *obj = root.get("MAS");
        if (obj == nullptr)
        {
            // Inner object given directly (extension); otherwise nothing to read, as the
            // reference's MASSolver::set_parameters returns when "MAS" is absent.
            bool any = false;
            for (const char *k : known_keys)
                any = any || root.get(k) != nullptr;
            if (!any)
                return;
            obj = &root;
        }
        if (obj->kind == JValue::Null)
            return;
        if (obj->kind != JValue::Object)
            throw InvalidArgumentError("[MAS] parameters[\"MAS\"] must be an object");

        Params q = p;
        if (const JValue *v = obj->get("block_dim"))
        {
            q.block_dim = as_int(*v, "block_dim");
            if (q.block_dim < 1 || q.block_dim > 3)
                bad("block_dim", "must be 1, 2 or 3");
        }
        if (const JValue *v = obj->get("max_iter"))
            q.max_iter = as_int(*v, "max_iter");
        if (const JValue *v = obj->get("relative_tolerance"))
            q.rel_tol = as_double(*v, "relative_tolerance");
        if (const JValue *v = obj->get("absolute_tolerance"))
            q.abs_tol = as_double(*v, "absolute_tolerance");
        if (const JValue *v = obj->get("lazy_partitioning"))
            q.lazy_partitioning = as_bool(*v, "lazy_partitioning");
        if (const JValue *v = obj->get("use_preconditioned_residual_norm"))
            q.use_preconditioned_residual_norm = as_bool(*v, "use_preconditioned_residual_norm");
        if (const JValue *v = obj->get("true_residual_period"))
        {
            q.true_residual_period = as_int(*v, "true_residual_period");
            if (q.true_residual_period < 1)
                bad("true_residual_period", "must be >= 1");
        }
        if (const JValue *v = obj->get("precision"))
        {
            const std::string s = as_string(*v, "precision");
            if (s == "fp64")
                q.precision = Precision::FP64;
            else if (s == "mixed")
                q.precision = Precision::Mixed;
            else
                bad("precision", "must be \"fp64\" or \"mixed\"");
        }
        if (const JValue *v = obj->get("aggregation"))
        {
            const std::string s = as_string(*v, "aggregation");
            if (s == "auto")
                q.aggregation = Aggregation::Auto;
            else if (s == "morton")
                q.aggregation = Aggregation::Morton;
            else if (s == "graph")
                q.aggregation = Aggregation::Graph;
            else
                bad("aggregation", "must be \"auto\", \"morton\" or \"graph\"");
        }
        if (const JValue *v = obj->get("coarse_solver"))
        {
            const std::string s = as_string(*v, "coarse_solver");
            if (s == "bank")
                q.coarse_solver = CoarseSolver::Bank;
            else if (s == "cholesky")
                q.coarse_solver = CoarseSolver::Cholesky;
            else
                bad("coarse_solver", "must be \"bank\" or \"cholesky\"");
        }
        if (const JValue *v = obj->get("reliable_update_delta"))
        {
            q.reliable_update_delta = as_double(*v, "reliable_update_delta");
            if (!(q.reliable_update_delta > 0.0 && q.reliable_update_delta < 1.0))
                bad("reliable_update_delta", "must be in (0, 1)");
        }
        if (const JValue *v = obj->get("check_interval"))
        {
            q.check_interval = as_int(*v, "check_interval");
            if (q.check_interval < 1 || q.check_interval > 1000)
                bad("check_interval", "must be in [1, 1000]");
        }
        if (const JValue *v = obj->get("use_graphs"))
            q.use_graphs = as_bool(*v, "use_graphs");
        if (const JValue *v = obj->get("learned_correction"))
            q.learned_correction = as_bool(*v, "learned_correction");
        if (const JValue *v = obj->get("division_guard"))
        {
            q.division_guard = as_double(*v, "division_guard");
            if (!(q.division_guard >= 0.0))
                bad("division_guard", "must be >= 0");
        }
        if (const JValue *v = obj->get("learned_collect"))
            q.learned_collect = as_bool(*v, "learned_collect");
        if (const JValue *v = obj->get("lattice_points"))
        {
            q.lattice_points = as_int(*v, "lattice_points");
            if (q.lattice_points < 2 || q.lattice_points > 16)
                bad("lattice_points", "must be in [2, 16]");
        }
        if (const JValue *v = obj->get("learned_hidden"))
        {
            q.learned_hidden = as_int(*v, "learned_hidden");
            if (q.learned_hidden < 8 || q.learned_hidden > 1024 || q.learned_hidden % 8 != 0)
                bad("learned_hidden", "must be a multiple of 8 in [8, 1024]");
        }
        if (const JValue *v = obj->get("learned_random_pairs"))
        {
            q.learned_random_pairs = as_int(*v, "learned_random_pairs");
            if (q.learned_random_pairs < 0 || q.learned_random_pairs > 100000)
                bad("learned_random_pairs", "must be in [0, 100000]");
        }
        if (const JValue *v = obj->get("learned_snapshots"))
        {
            q.learned_snapshots = as_int(*v, "learned_snapshots");
            if (q.learned_snapshots < 0 || q.learned_snapshots > 1000)
                bad("learned_snapshots", "must be in [0, 1000]");
        }
        if (const JValue *v = obj->get("spmv_unroll"))
        {
            q.spmv_unroll = as_int(*v, "spmv_unroll");
            if (q.spmv_unroll != 2 && q.spmv_unroll != 4 && q.spmv_unroll != 8)
                bad("spmv_unroll", "must be 2, 4 or 8");
        }
        if (const JValue *v = obj->get("spmv_halo"))
            q.spmv_halo = as_bool(*v, "spmv_halo");
        if (const JValue *v = obj->get("device_setup"))
            q.device_setup = as_bool(*v, "device_setup");
        if (const JValue *v = obj->get("spmv_copies"))
            q.spmv_copies = as_bool(*v, "spmv_copies");
        if (const JValue *v = obj->get("spmv_fuse_direction"))
            q.spmv_fuse_direction = as_bool(*v, "spmv_fuse_direction");
        if (const JValue *v = obj->get("restrict_piece"))
        {
            q.restrict_piece = as_int(*v, "restrict_piece");
            if (q.restrict_piece <= RESTRICT_SMALL_MAX || q.restrict_piece > RESTRICT_PIECE)
                bad("restrict_piece", "must be in [9, 512]");
        }
        if (const JValue *v = obj->get("learned_network"))
        {
            const std::string k = as_string(*v, "learned_network");
            if (k != "mlp" && k != "linear")
                bad("learned_network", "must be \"mlp\" or \"linear\"");
            q.learned_linear = k == "linear";
        }
        if (const JValue *v = obj->get("apply_wave"))
            q.apply_wave = as_bool(*v, "apply_wave");
        if (const JValue *v = obj->get("single_reduction"))
            q.single_reduction = as_bool(*v, "single_reduction");
        if (const JValue *v = obj->get("spmv_symmetric"))
            q.spmv_symmetric = as_bool(*v, "spmv_symmetric");
        if (const JValue *v = obj->get("apply_workgroups"))
        {
            q.apply_workgroups = as_int(*v, "apply_workgroups");
            if (q.apply_workgroups < 0 || q.apply_workgroups > 1000000)
                bad("apply_workgroups", "must be in [0, 1000000]");
        }
        if (const JValue *v = obj->get("learned_output_scale"))
        {
            q.learned_output_scale = as_double(*v, "learned_output_scale");
            if (!std::isfinite(q.learned_output_scale))
                bad("learned_output_scale", "must be finite");
        }
        if (const JValue *v = obj->get("learned_stall_segments"))
        {
            q.learned_stall_segments = as_int(*v, "learned_stall_segments");
            if (q.learned_stall_segments < 1 || q.learned_stall_segments > 1000)
                bad("learned_stall_segments", "must be in [1, 1000]");
        }
        p = q;
    }

    std::string params_to_json(const Params &p)
    {
        std::ostringstream o;
        o.precision(17);
        o << "{\"block_dim\": " << p.block_dim << ", \"max_iter\": " << p.max_iter
          << ", \"relative_tolerance\": " << json_number(p.rel_tol) << ", \"absolute_tolerance\": " << json_number(p.abs_tol)
          << ", \"lazy_partitioning\": " << (p.lazy_partitioning ? "true" : "false")
          << ", \"use_preconditioned_residual_norm\": " << (p.use_preconditioned_residual_norm ? "true" : "false")
          << ", \"true_residual_period\": " << p.true_residual_period
          << ", \"precision\": \"" << (p.precision == Precision::FP64 ? "fp64" : "mixed") << "\""
          << ", \"aggregation\": \""
          << (p.aggregation == Aggregation::Auto ? "auto" : (p.aggregation == Aggregation::Morton ? "morton" : "graph"))
          << "\", \"coarse_solver\": \"" << (p.coarse_solver == CoarseSolver::Bank ? "bank" : "cholesky") << "\""
          << ", \"reliable_update_delta\": " << json_number(p.reliable_update_delta)
          << ", \"check_interval\": " << p.check_interval
          << ", \"use_graphs\": " << (p.use_graphs ? "true" : "false")
          << ", \"learned_correction\": " << (p.learned_correction ? "true" : "false")
          << ", \"division_guard\": " << json_number(p.division_guard)
          << ", \"learned_collect\": " << (p.learned_collect ? "true" : "false")
          << ", \"lattice_points\": " << p.lattice_points << ", \"learned_hidden\": " << p.learned_hidden
          << ", \"learned_random_pairs\": " << p.learned_random_pairs
          << ", \"learned_snapshots\": " << p.learned_snapshots
          << ", \"learned_stall_segments\": " << p.learned_stall_segments
          << ", \"learned_output_scale\": " << json_number(p.learned_output_scale) << ", \"spmv_unroll\": " << p.spmv_unroll << ", \"spmv_halo\": " << (p.spmv_halo ? "true" : "false") << ", \"device_setup\": " << (p.device_setup ? "true" : "false") << ", \"spmv_copies\": " << (p.spmv_copies ? "true" : "false") << ", \"spmv_fuse_direction\": " << (p.spmv_fuse_direction ? "true" : "false") << ", \"restrict_piece\": " << p.restrict_piece
          << ", \"apply_workgroups\": " << p.apply_workgroups << ", \"apply_wave\": " << (p.apply_wave ? "true" : "false")
          << ", \"single_reduction\": " << (p.single_reduction ? "true" : "false")
          << ", \"spmv_symmetric\": " << (p.spmv_symmetric ? "true" : "false")
          << ", \"learned_network\": \"" << (p.learned_linear ? "linear" : "mlp") << "\"}";
        return o.str();
    }
} // namespace psmas
