// XmlLogger.h -- the reference's run log (src/XmlLogger.h:8-93): an XML file
// <logName>_log.xml of <Cycle ID="..."> elements holding named <Section>s of
// scalar tags, and <logName>_timeseries.csv with one probe row per logged
// cycle.  Sections and tags are the reference's, so its analysis scripts
// (tools/parseXML.py, elapsedTime.py, totalEnergy.py) read the log as they
// read the reference's.
#pragma once
#include <fstream>
#include <map>
#include <string>

namespace gample {

class XmlLogger {
public:
    XmlLogger(const std::string &logName, const std::string &csvHeader);
    ~XmlLogger();
    XmlLogger(const XmlLogger &) = delete;
    XmlLogger &operator=(const XmlLogger &) = delete;
    // sections and CSV cells are written only while the switch is on (the
    // driver turns it on for the cycles that are multiples of LogOutput)
    void switchLog(bool on) { on_ = on; }
    bool on() const { return on_; }
    void logCycleStart(int cycle, double time);
    void logSection(const std::string &name, const std::map<std::string, std::string> &data);
    void csvOutput(const std::string &cells);
    void logCycleEnd();
    // a CDATA block, written whatever the switch
    void logComment(const std::string &text);

private:
    std::ofstream xml_, csv_;
    bool on_ = false;
};

// v as a float in scientific notation with prec digits (the reference's
// fmtSci takes a float, so doubles are rounded to float first)
std::string fmtSci(float v, int prec);

}  // namespace gample
