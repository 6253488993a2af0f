// nnp::Log (Log.h of the reference): log entries kept in memory and echoed
// to stdout and to every registered C file pointer and output stream.
#ifndef NNP_LOG_H
#define NNP_LOG_H

#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

namespace nnp
{

class Log
{
public:
    Log() : writeToStdout(true), silent(false) {}
    Log& operator<<(std::string const& entry);
    Log& operator<<(std::vector<std::string> const& entries);
    void addLogEntry(std::string const& entry);
    void addMultipleLogEntries(std::vector<std::string> const& entries);
    void registerCFilePointer(FILE** const& filePointer);
    void registerStreamPointer(std::ofstream* const& streamPointer);
    std::vector<std::string> getLog() const { return memory; }

    bool writeToStdout;
    bool silent;

private:
    std::vector<std::string> memory;
    std::vector<FILE**> cFilePointers;
    std::vector<std::ofstream*> streamPointers;
};

}  // namespace nnp

#endif
