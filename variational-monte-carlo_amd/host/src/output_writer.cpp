// output_writer.cpp -- the three frame writers (formats fixed by the
// reference's src/output_writer/output_writer.cpp and its consumer render.py).
#include "output_writer/output_writer.hpp"

#include <limits>
#include <stdexcept>
#include <string_view>

namespace {

// JSON string literal with the escapes of RFC 8259 (control characters as
// \u00XX, the short forms where JSON has one).
void json_string(std::ostream &out, std::string_view text) {
    static constexpr char hex[] = "0123456789abcdef";
    out << '"';
    for (char c : text) {
        const auto u = static_cast<unsigned char>(c);
        switch (c) {
            case '"': out << "\\\""; break;
            case '\\': out << "\\\\"; break;
            case '\b': out << "\\b"; break;
            case '\f': out << "\\f"; break;
            case '\n': out << "\\n"; break;
            case '\r': out << "\\r"; break;
            case '\t': out << "\\t"; break;
            default:
                if (u < 0x20U) out << "\\u00" << hex[u >> 4] << hex[u & 15U];
                else out << c;
        }
    }
    out << '"';
}

void json_numbers(std::ostream &out, const std::vector<double> &values) {
    out << '[';
    const char *sep = "";
    for (double v : values) {
        out << sep << v;
        sep = ",";
    }
    out << ']';
}

template <typename T>
void put(std::ostream &out, const T &v) {
    out.write(reinterpret_cast<const char *>(&v), sizeof(T));
}

}  // namespace

JsonOutputWriter::JsonOutputWriter(std::ostream &out) : out_{out} {
    out_.precision(std::numeric_limits<double>::max_digits10);
}

void JsonOutputWriter::write_init(const InitData &d) {
    out_ << "{\"type\":\"init\",\"runId\":";
    json_string(out_, d.run_id);
    out_ << ",\"numParticles\":" << d.num_particles << ",\"boxLength\":" << d.box_length
         << ",\"warmupSteps\":" << d.warmup_steps << ",\"measureSteps\":" << d.measure_steps
         << ",\"stepSize\":" << d.step_size << ",\"blockSize\":" << d.block_size
         << ",\"seed\":" << d.seed << "}\n";
}

void JsonOutputWriter::write_frame(const FrameData &d) {
    out_ << "{\"type\":\"frame\",\"step\":" << d.step << ",\"accepted\":" << d.accepted
         << ",\"proposed\":" << d.proposed << ",\"acceptanceRate\":" << d.acceptance_rate
         << ",\"localEnergy\":" << d.local_energy << ",\"meanEnergy\":" << d.mean_energy;
    if (d.standard_error) out_ << ",\"standardErrorAvailable\":true,\"standardError\":" << *d.standard_error;
    else out_ << ",\"standardErrorAvailable\":false,\"standardError\":0";
    out_ << ",\"positions\":";
    json_numbers(out_, d.positions);
    out_ << "}\n";
}

void JsonOutputWriter::write_done(const DoneData &d) {
    out_ << "{\"type\":\"done\",\"totalAccepted\":" << d.total_accepted
         << ",\"totalProposed\":" << d.total_proposed
         << ",\"finalAcceptanceRate\":" << d.final_acceptance_rate
         << ",\"finalMeanEnergy\":" << d.final_mean_energy;
    if (d.final_standard_error)
        out_ << ",\"finalStandardErrorAvailable\":true,\"finalStandardError\":" << *d.final_standard_error;
    else out_ << ",\"finalStandardErrorAvailable\":false,\"finalStandardError\":0";
    out_ << "}\n";
}

CsvOutputWriter::CsvOutputWriter(std::ostream &out) : out_{out} {}
void CsvOutputWriter::write_init(const InitData &) { throw std::logic_error{"CsvOutputWriter is not implemented yet"}; }
void CsvOutputWriter::write_frame(const FrameData &) { throw std::logic_error{"CsvOutputWriter is not implemented yet"}; }
void CsvOutputWriter::write_done(const DoneData &) { throw std::logic_error{"CsvOutputWriter is not implemented yet"}; }

BinOutputWriter::BinOutputWriter(std::ostream &out) : out_{out} {}

void BinOutputWriter::write_init(const InitData &d) {
    put(out_, static_cast<uint64_t>(d.num_particles));
    put(out_, d.box_length);
    put(out_, static_cast<uint64_t>(d.measure_steps));
    out_.flush();
}

void BinOutputWriter::write_frame(const FrameData &d) {
    put(out_, d.local_energy);
    put(out_, d.mean_energy);
    put(out_, d.standard_error.value_or(0.0));
    put(out_, d.acceptance_rate);
    out_.write(reinterpret_cast<const char *>(d.positions.data()),
               static_cast<std::streamsize>(d.positions.size() * sizeof(double)));
}

void BinOutputWriter::write_done(const DoneData &) { out_.flush(); }

std::unique_ptr<OutputWriter> make_output_writer(OutputFormat format, std::ostream &out) {
    switch (format) {
        case OutputFormat::JSON: return std::make_unique<JsonOutputWriter>(out);
        case OutputFormat::CSV: return std::make_unique<CsvOutputWriter>(out);
        case OutputFormat::BIN: return std::make_unique<BinOutputWriter>(out);
    }
    throw std::invalid_argument("Unsupported output format");
}
