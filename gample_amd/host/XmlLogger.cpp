// XmlLogger.cpp -- see XmlLogger.h.
#include "XmlLogger.h"

#include <cstdio>
#include <stdexcept>

namespace gample {

XmlLogger::XmlLogger(const std::string &logName, const std::string &csvHeader)
{
    xml_.open(logName + "_log.xml");
    if (!xml_) throw std::runtime_error("XmlLogger: " + logName + "_log.xml could not be created");
    xml_ << "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n<SimulationLog>\n";
    csv_.open(logName + "_timeseries.csv");
    if (!csv_) throw std::runtime_error("XmlLogger: " + logName + "_timeseries.csv could not be created");
    csv_ << csvHeader << "\n";
}

XmlLogger::~XmlLogger() { xml_ << "</SimulationLog>\n"; }

void XmlLogger::logCycleStart(int cycle, double time)
{
    if (!on_) return;
    // both streams print the time with the stream default (6 significant digits)
    xml_ << " <Cycle ID=\"" << cycle << "\">\n  time = " << time << "\n";
    csv_ << cycle << "," << time;
}

void XmlLogger::logSection(const std::string &name, const std::map<std::string, std::string> &data)
{
    if (!on_) return;
    xml_ << "  <Section Name=\"" << name << "\">\n";
    for (const auto &kv : data) xml_ << "   <" << kv.first << ">" << kv.second << "</" << kv.first << ">\n";
    xml_ << "  </Section>\n";
}

void XmlLogger::csvOutput(const std::string &cells)
{
    if (on_) csv_ << "," << cells;
}

void XmlLogger::logCycleEnd()
{
    if (!on_) return;
    xml_ << " </Cycle>\n";
    csv_ << "\n";
    xml_.flush();
    csv_.flush();
}

void XmlLogger::logComment(const std::string &text) { xml_ << "<![CDATA[\n" << text << "]]>\n"; }

// d.ddde+XX with prec digits after the point (the log's number format)
std::string fmtSci(float v, int prec)
{
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.*e", prec, (double)v);
    return std::string(buf);
}

}  // namespace gample
