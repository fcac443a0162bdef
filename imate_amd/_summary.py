"""``verbose=True`` report: the results / config / process tables of imate's
print_summary (imate/_trace_estimator/trace_estimator_utilities.py:380-571),
same fields and layout."""

import numpy


def _fmt_val(x, ok, bad):
    return (bad % x) if (numpy.isnan(x) or numpy.isinf(x)) else (ok % x)


def print_summary(info):
    m, e, c, t, d, s = (info['matrix'], info['error'], info['convergence'],
                        info['time'], info['device'], info['solver'])
    data_type = m['data_type'].decode('utf-8')
    bits = {'float32': '32-bit', 'float64': '64-bit',
            'float128': '128-bit'}[data_type]
    q = m['num_inquiries']
    npar = m['num_operator_parameters']
    ortho = s['orthogonalize']
    if ortho < 0 or ortho >= s['lanczos_degree']:
        ortho = 'full'
    elif ortho == 0:
        ortho = 'none'
    else:
        ortho = '%4d' % ortho
    rule = '=' * 78
    print('                                    results' + ' ' * 28)
    print(rule)
    print('     inquiries                            error            samples')
    print('--------------------              ---------------------   ---------')
    print('i         parameters       trace    absolute   relative   num   out'
          '  converged')
    print(rule)
    trace = numpy.atleast_1d(c['samples_mean'])
    aerr = numpy.atleast_1d(e['absolute_error'])
    rerr = numpy.atleast_1d(e['relative_error']) * 100.0
    used = numpy.atleast_1d(c['num_samples_used'])
    outl = numpy.atleast_1d(c['num_outliers'])
    conv = numpy.atleast_1d(c['converged'])
    for i in range(q):
        line = '%-3d  ' % (i + 1)
        if npar == 0:
            line += ' %8s  none  ' % ''
        elif npar == 1:
            line += '%15.4e  ' % m['parameters'][i]
        else:
            line += '[%7.2e, ...]  ' % m['parameters'][i * npar]
        line += _fmt_val(trace[i], '%+7.3e   ', '%10e   ')
        line += _fmt_val(aerr[i], '%7.3e  ', '%9e  ')
        line += _fmt_val(rerr[i], '%8.3f%%  ', '%9f  ')
        line += '%4d  %4d' % (used[i], outl[i])
        line += '%11s' % bool(conv[i])
        print(line)
    print('')
    print('                                    config' + ' ' * 29)
    print(rule)
    print('                matrix                            '
          'stochastic estimator        ')
    print('-------------------------------------    '
          '-------------------------------------')
    print('gram:                           %5s' % m['gram'], end='    ')
    print('method:                           slq')
    ex = m['exponent']
    if int(ex) == ex:
        print('exponent:                         %3d' % int(ex), end='    ')
    else:
        print('exponent:                      %6.2f' % ex, end='    ')
    print('lanczos degree:                  %4d' % s['lanczos_degree'])
    print('num matrix parameters:             %2d' % npar, end='    ')
    print('lanczos tol:                %8.3e' % s['lanczos_tol'])
    print('data type:                   %8s' % bits, end='    ')
    print('orthogonalization:               %4s' % ortho)
    print('')
    print('             convergence                                 error'
          '                ')
    print('-------------------------------------    '
          '-------------------------------------')
    print('min num samples:                %5d' % c['min_num_samples'],
          end='    ')
    print('abs error tol:              %8.3e' % e['error_atol'])
    print('max num samples:                %5d' % c['max_num_samples'],
          end='    ')
    print('rel error tol:                  %4.2f%%' % (e['error_rtol'] * 100.0))
    print('outlier significance level:     %4.2f%%'
          % e['outlier_significance_level'], end='    ')
    print('confidence level:              %4.2f%%'
          % (e['confidence_level'] * 100.0))
    print('')
    print('                                   process' + ' ' * 30)
    print(rule)
    print('                 time                                   device'
          '                ')
    print('-------------------------------------    '
          '-------------------------------------')
    print('tot wall time (sec):        %8.3e' % t['tot_wall_time'], end='    ')
    print('num cpu threads:                  %3d' % d['num_cpu_threads'])
    print('alg wall time (sec):        %8.3e' % t['alg_wall_time'], end='    ')
    print('num gpu devices, multiproc:    %2d, %2d'
          % (d['num_gpu_devices'], d['num_gpu_multiprocessors']))
    print('cpu proc time (sec):        %8.3e' % t['cpu_proc_time'], end='    ')
    print('num gpu threads per multiproc:   %4d'
          % d['num_gpu_threads_per_multiprocessor'])
    print('')
